/*
 * ytrc.h -- C ABI of libytrc.so, the MI355X (gfx950) brick ray-caster.
 *
 * The reference's boundary for this path is the ImageSampler protocol of
 * yt/utilities/lib/image_samplers.pyx:
 *   ImageSampler(vp_pos, vp_dir, center, bounds, image, x_vec, y_vec, width,
 *                volume_method, ...)            image_samplers.pyx:59-130
 *   sampler(pg: PartitionedGrid, num_threads)   image_samplers.pyx:132-197
 *   sampler.aimage                              image_samplers.pxd:52
 * with bricks described by PartitionedGrid / VolumeContainer
 * (partitioned_grid.pyx:28-63, volume_container.pxd:13-33).
 *
 * This ABI splits that protocol into: a device-resident dataset of bricks
 * (the PartitionedGrid data the sampler is called with), a request (the
 * sampler constructor arguments plus the brick call order), and a batch of
 * requests that runs on the GPU and is read back into the caller's image.
 *
 * All functions return 0 on success and a negative YTRC_E* code on error;
 * ytrc_last_error() returns a thread-local message for the last error.
 * Plain pointers and sizes only.  Host pointers unless stated otherwise.
 */
#ifndef YTRC_H
#define YTRC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YTRC_OK 0
#define YTRC_EINVAL (-1)    /* bad argument (the reference raises ValueError/RuntimeError) */
#define YTRC_EDEVICE (-2)   /* HIP runtime error                                          */
#define YTRC_EBUDGET (-3)   /* request would exceed the context's HBM budget              */
#define YTRC_ENOTIMPL (-4)  /* e.g. unknown volume method (image_samplers.pyx:137-141)    */
#define YTRC_ECANCELLED (-5)
#define YTRC_EBUSY (-6)     /* the dataset is still used by a batch or a server request    */
#define YTRC_ETIMEDOUT (-7) /* a wait's deadline passed before the work finished           */

/* Sampler kinds: ProjectionSampler (image_samplers.pyx:327-343),
 * VolumeRenderSampler (image_samplers.pyx:405-494) and
 * InterpolatedProjectionSampler (image_samplers.pyx:344-402, the sampler of
 * a "projection" VolumeSource, render_source.py:464-488; vertex-centred
 * bricks, n_samples trilinear sub-samples per cell, min(3, n_fields)
 * channels). */
#define YTRC_PROJECTION 0
#define YTRC_VOLUME_RENDER 1
#define YTRC_INTERPOLATED_PROJECTION 2
/* LightSourceRenderSampler (image_samplers.pyx:496-600): the volume render's
 * tables with a directional light (light_dir, light_rgba of the request),
 * no mask test (the sampler of old_camera.py:688). */
#define YTRC_LIGHT_SOURCE 3

/* Lenses: plane-parallel (lenses.pyx:22-55, :179-192), and per-pixel rays
 * (generate_vector_info_null, lenses.pyx:196-207), which is how yt runs the
 * perspective, stereo, fisheye and spherical lenses (lens.py:145-800): the
 * lens computes an origin and a direction per pixel on the host. */
#define YTRC_LENS_PLANE_PARALLEL 0
#define YTRC_LENS_PER_PIXEL 1

/* Request flags. */
#define YTRC_FLAG_RECORD 1u /* per-pixel visit count, visit hash and sample count */
#define YTRC_FLAG_ERT 2u    /* early ray termination (VR only; changes op order)  */
/* VolumeRenderSampler only, plane-parallel lens, no image_in: instead of an
 * image, the request yields its per-pixel, per-channel affine map
 * rgba -> C + T rgba over its bricks (read with ytrc_batch_read_partial).
 * The maps of consecutive runs of a brick order compose to the map of the
 * whole order, for every transfer function (grey or not). */
#define YTRC_FLAG_PARTIAL 4u

/* Batch flags. */
#define YTRC_BATCH_PLAIN 0u      /* one kernel launch per request                    */
#define YTRC_BATCH_PERSISTENT 1u /* one persistent launch draining a global tile queue */
#define YTRC_BATCH_GRAPH 2u      /* capture the launch sequence into a hipGraph       */
/* Rays per 64-lane tile (rows along image axis 0 x columns along axis 1). */
#define YTRC_BATCH_TILE_8x8 0u
#define YTRC_BATCH_TILE_4x16 16u
#define YTRC_BATCH_TILE_16x4 32u
/* Tiles per work item of a persistent launch (rays per item: 64 x this):
 * one wave renders an item's consecutive tiles of one image in a row. */
#define YTRC_BATCH_ITEM_TILES_1 0u
#define YTRC_BATCH_ITEM_TILES_2 64u
#define YTRC_BATCH_ITEM_TILES_4 128u
#define YTRC_BATCH_ITEM_TILES_8 192u
/* Transfer-function table bytes one persistent launch stages in each
 * workgroup's LDS: a launch holds the distinct tables of its requests; with
 * a cap, a kernel variant's requests are split over launches by table (a
 * table is never split; one larger than the cap gets a launch of its own).
 * Fewer LDS bytes per workgroup let more workgroups share a CU.  Images do
 * not depend on it. */
#define YTRC_BATCH_TF_LDS_ALL 0u   /* one launch per kernel variant          */
#define YTRC_BATCH_TF_LDS_32K 256u /* at most 32 KiB of tables per launch    */
#define YTRC_BATCH_TF_LDS_16K 512u /* at most 16 KiB of tables per launch    */
#define YTRC_BATCH_TF_LDS_ONE 768u /* one distinct table per launch          */
/* Persistent batches: volume renders with a ColorTransferFunction, one
 * field and 10 sub-samples stage vertex blocks in LDS, 512 bytes per wave:
 * at a visit where all lanes of a wave are in one block, the wave reads it
 * once and takes the corners from LDS (the L2 path otherwise).  Images do
 * not depend on it. */
#define YTRC_BATCH_VERTEX_LDS 1024u

typedef struct ytrc_ctx ytrc_ctx;
typedef struct ytrc_batch ytrc_batch;

/* Context: one per device.  budget_bytes caps the device memory this context
 * may hold (datasets + batches); 0 means no cap. */
int ytrc_ctx_create(int32_t device, uint64_t budget_bytes, ytrc_ctx **out);
void ytrc_ctx_destroy(ytrc_ctx *ctx);
int ytrc_ctx_used_bytes(const ytrc_ctx *ctx, uint64_t *used, uint64_t *budget);
/* LDS bytes one workgroup may allocate on the context's device (HIP's
 * sharedMemPerBlock): it bounds the transfer-function tables of one launch,
 * which sit in LDS (the reference keeps them in host memory,
 * field_interpolation_tables.pxd:27-50, and has no such bound). */
int ytrc_ctx_max_lds(const ytrc_ctx *ctx, uint64_t *bytes);
/* The device's name, ISA and CU count as one NUL-terminated string, cut to
 * cap bytes (including the NUL).  Labels tuning caches and
 * reports (the reference runs on the host and has no counterpart). */
int ytrc_ctx_device_name(const ytrc_ctx *ctx, char *buf, int32_t cap);
const char *ytrc_last_error(void);

/* Brick descriptor: PartitionedGrid(parent_grid_id, data, mask, left_edge,
 * right_edge, dims) (partitioned_grid.pyx:28-63).  vertex_centred = 1 when
 * each field holds (dims+1)^3 vertex values (AMRKDTree bricks for
 * VolumeRenderSampler, amr_kdtree.py:360-367); 0 for dims^3 cell values
 * (off_axis_projection bricks, off_axis_projection.py:506-522). */
typedef struct {
    double left_edge[3];
    double right_edge[3];
    int32_t dims[3];
    int32_t vertex_centred;
} ytrc_brick_desc;

/* Copies n_bricks bricks to the device.  field_ptrs holds n_bricks*n_fields
 * host pointers, brick-major, each a C-order array of src_dtype
 * (0 = float64, 1 = float32) with dims or dims+1 points per axis.
 * mask_ptrs: NULL, or n_bricks pointers to dims-shaped uint8 masks (a NULL
 * entry means all ones).  Device data is float32, fields interleaved per
 * point; cell-centred bricks in blocks of 4x4x4 cells, vertex-centred ones
 * in blocks of the 5x5x5 vertices of 4x4x4 cells (face vertices stored in
 * both blocks).  Replaces the per-call numpy views of PartitionedGrid. */
int ytrc_dataset_upload(ytrc_ctx *ctx, int32_t n_bricks, const ytrc_brick_desc *bricks,
                        int32_t n_fields, int32_t src_dtype, const void *const *field_ptrs,
                        const uint8_t *const *mask_ptrs, int32_t *ds_id_out);
/* Device brick preparation for VolumeRenderSampler: AMRKDTree.get_brick_data
 * (amr_kdtree.py:325-374) run on the GPU.  The caller passes each grid's
 * cell-centred data once (n_grids*n_fields host pointers, grid-major, C-order
 * dims^3 arrays of src_dtype) and the kd leaves (the grid each lies in, its
 * cell slice [li, ri) and its edges); the library computes every leaf's
 * vertex data with no-ghost vertex centring (grid_patch.py:270-296,
 * interpolators.pyx:177-226) in FP64, applies log10 with v <= 0 -> NaN to
 * the fields flagged in log_fields (amr_kdtree.py:346-352), and stores the
 * leaves as vertex-centred bricks, in leaf order, exactly as
 * ytrc_dataset_upload would store the host-computed bricks.  mask_ptrs:
 * NULL or n_leaves child-mask slices (dims^3 uint8, NULL = all ones).
 * Grids need at least 2 cells per axis. */
typedef struct {
    double left_edge[3];
    double right_edge[3];
    int32_t dims[3];
    int32_t pad;
} ytrc_grid_desc;
typedef struct {
    int32_t grid;
    int32_t li[3];
    int32_t ri[3];
    int32_t pad;
    double left_edge[3];
    double right_edge[3];
} ytrc_leaf_desc;
int ytrc_dataset_build_vr(ytrc_ctx *ctx, int32_t n_grids, const ytrc_grid_desc *grids,
                          int32_t n_fields, int32_t src_dtype, const void *const *cell_ptrs,
                          const int32_t *log_fields, int32_t n_leaves, const ytrc_leaf_desc *leaves,
                          const uint8_t *const *mask_ptrs, int32_t *ds_id_out);
/* Copies brick `brick`'s stored points (float32, fields interleaved, in
 * storage order) to out; *n_out gets the float count (out may be NULL to
 * query it).  No counterpart in the reference: used to check the device
 * brick preparation against the host's. */
int ytrc_dataset_read_points(ytrc_ctx *ctx, int32_t ds_id, int32_t brick, float *out, int64_t cap,
                             int64_t *n_out);
/* Frees a dataset's device memory.  Refused with YTRC_EBUSY while a live
 * batch or an unread server request uses it (a batch holds its datasets
 * until ytrc_batch_destroy; a server request until it is read, cancelled or
 * the server is destroyed). */
int ytrc_dataset_free(ytrc_ctx *ctx, int32_t ds_id);
int ytrc_dataset_bytes(const ytrc_ctx *ctx, int32_t ds_id, uint64_t *bytes);

/* Transfer function: the FieldInterpolationTable set a VolumeRenderSampler
 * builds from tf_obj (image_samplers.pyx:425-447,
 * field_interpolation_tables.pxd:27-61).  At most 6 tables. */
typedef struct {
    int32_t n_fits;
    int32_t nbins[6];
    double bounds[12]; /* (x_bounds[0], x_bounds[1]) per table */
    const double *y[6];
    int32_t field_ids[6];
    int32_t weight_field_ids[6];
    int32_t weight_table_ids[6];
    int32_t field_table_ids[6];
    int32_t grey_opacity;
} ytrc_tf_desc;

/* One image: the ImageSampler constructor arguments (image_samplers.pyx:59-130,
 * built by utils.py:55-139 from PlaneParallelLens._get_sampler_params,
 * lens.py:83-112) plus the brick call order (the sequence of sampler(pg)
 * calls, render_source.py:594-597 / off_axis_projection.py:506-524). */
typedef struct {
    int32_t kind;
    int32_t lens;
    int32_t nv[2];
    double vp_pos[12];
    double vp_dir[3];
    double center[3];
    double bounds[4];
    double x_vec[3];
    double y_vec[3];
    double width[3];
    int32_t n_samples;           /* VR sub-samples per cell (render_source.py:246) */
    int32_t ds_id;
    int32_t n_order;
    const int32_t *order;        /* brick indices into the dataset, call order    */
    const double *zbuffer;       /* nv0*nv1 or NULL (ones, utils.py:136)           */
    const double *image_in;      /* nv0*nv1*4 initial image or NULL (zeros)        */
    const ytrc_tf_desc *tf;      /* VR only                                        */
    uint32_t flags;
    /* YTRC_LENS_PER_PIXEL only: ray origins and directions, nv0*nv1*3
     * doubles each (the sampler's vp_pos[vi, vj, :] and vp_dir[vi, vj, :]);
     * camera_data: NULL (every brick covers the whole image,
     * calculate_extent_null, lenses.pyx:169-175) or the perspective
     * camera's 5x3 rows position, width, east, north, normal (camera.py:280)
     * for calculate_extent_perspective (lenses.pyx:60-161,
     * image_samplers.pyx:98-99). */
    const double *pix_pos;
    const double *pix_dir;
    const double *camera_data;
    /* YTRC_LIGHT_SOURCE only: unit light direction and light colour
     * (image_samplers.pyx:510-531 normalises the direction). */
    double light_dir[3];
    double light_rgba[4];
} ytrc_request;

/* Builds a batch: validates requests, computes the per-brick pixel extents
 * (lenses.pyx:22-55 with the +-1 clip of image_samplers.pyx:159-166) and
 * stages every descriptor and output buffer in device memory. */
int ytrc_batch_create(ytrc_ctx *ctx, int32_t n_req, const ytrc_request *reqs,
                      uint32_t batch_flags, ytrc_batch **out);
/* Enqueues the batch's kernels on the context stream; returns immediately.
 * A persistent batch of several launches runs all but the heaviest on the
 * context's second stream, joined back into the first before the launch's
 * end event. */
int ytrc_batch_launch(ytrc_batch *b);
/* Waits for the last launch; ms_out (may be NULL) gets the GPU time between
 * the HIP events recorded around it on the same stream. */
int ytrc_batch_sync(ytrc_batch *b, float *ms_out);
/* ytrc_batch_sync with a deadline: YTRC_ETIMEDOUT when the launch is still
 * running after timeout_ms (it keeps running; sync again or cancel). */
int ytrc_batch_sync_for(ytrc_batch *b, double timeout_ms, float *ms_out);
/* Asks the running launch to stop, the analog of the reference's
 * PyErr_CheckSignals every 1000 rays (image_samplers.pyx:188-190): each
 * wave of a persistent launch (YTRC_BATCH_PERSISTENT) reads the batch's stop
 * flag (host-pinned) at every 16th work item it takes (each wave at its own
 * phase), and exits when it is set; plain launches finish the image they
 * render.  The
 * next sync returns YTRC_ECANCELLED and the images stay unreadable until a
 * later launch completes.  Safe to call from another thread while one
 * waits in ytrc_batch_sync. */
int ytrc_batch_cancel(ytrc_batch *b);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): renders one request with the persistent tile-queue
 * kernels, exactly as a YTRC_BATCH_PERSISTENT batch of that request does,
 * and copies its image to image_out (float64, nv0*nv1*4, as
 * ytrc_batch_read_image), calling no hipMalloc/hipFree.  Every staging buffer
 * (descriptor, brick order, extents, z-buffer, input image, output planes,
 * transfer-function table, tile list, queue counters) lives in one fixed
 * device array of YTRC_NOALLOC_SCRATCH_BYTES that the library holds on each
 * device until the process ends.  A request that needs more is refused
 * (YTRC_EINVAL), as are record and partial-map requests.  Synchronous.
 * Calls on one device share its array, which keeps the last call's contents;
 * the library serialises them (a call waits for the one before it). */
#define YTRC_NOALLOC_SCRATCH_BYTES (64u << 20)
int ytrc_render_noalloc(ytrc_ctx *ctx, const ytrc_request *r, double *image_out);
/* Copies request i's image to image_out (float64, nv0*nv1*4, the reference's
 * aimage layout, which cast_through_kdtree updates in float64,
 * image_samplers.pyx:177-196); channels the sampler does not write keep
 * image_in.  With an image_in, projection samplers' sums (computed from 0 on
 * the device) are added to it in float64, and compositing samplers' results
 * that still equal float(image_in) return image_in itself: a pixel no sample
 * changes keeps its input bits. */
int ytrc_batch_read_image(ytrc_batch *b, int32_t i, double *image_out);
/* YTRC_FLAG_PARTIAL only: request i's map to map_out (float64,
 * nv0*nv1*8: C[0..3] then T[0..3] per pixel).  Applied in call order, the
 * maps of the parts of a split brick order give the image of the whole
 * order.  Replaces the image exchange of yt's parallel volume render,
 * receive_and_reduce (yt/utilities/amr_kdtree/amr_kdtools.py:6-39, driven by
 * AMRKDTree.reduce_tree_images, amr_kdtree.py:297-323), which composites
 * images with the front alpha (grey opacity) or adds them (otherwise). */
int ytrc_batch_read_partial(ytrc_batch *b, int32_t i, double *map_out);
/* Record mode only: per-pixel visit count, visit hash, sample count. */
int ytrc_batch_read_record(ytrc_batch *b, int32_t i, int64_t *visits, uint64_t *hashes,
                           int64_t *samples);
/* Launch shape: waves (64-ray tiles) per workgroup (1, 2 or 4) and, for
 * persistent batches, workgroups per CU (0 = the occupancy limit). */
int ytrc_batch_set_launch(ytrc_batch *b, int32_t waves_per_group, int32_t groups_per_cu);
/* Plain, uncaptured batches: GPU time of request i's kernel in the last
 * launch, from HIP events recorded around it on the launch stream. */
int ytrc_batch_request_ms(ytrc_batch *b, int32_t i, float *ms);
/* Device bytes held by the batch (descriptors, extents, outputs). */
int ytrc_batch_bytes(const ytrc_batch *b, uint64_t *bytes);
/* Persistent batches: the kernel launches one launch of the batch makes
 * (n_launches) and the largest transfer-function table bytes one of them
 * stages in LDS (max_tf_bytes); set by the batch flags (YTRC_BATCH_TF_LDS_*).
 * No counterpart in the reference, which has no launches. */
int ytrc_batch_launch_plan(const ytrc_batch *b, int32_t *n_launches, uint64_t *max_tf_bytes);
void ytrc_batch_destroy(ytrc_batch *b);

/* Serving: continuous batching against resident datasets.  The reference
 * renders one sampler at a time with no queue (render_source.py:594-599,
 * off_axis_projection.py:506-526); a server admits requests while others
 * are on the GPU and runs the 8x8 ray tiles of every admitted request, in
 * arrival order (see ytrc_server_set_policy), in chunks of `chunk_tiles`
 * persistent-kernel work items
 * (up to two chunks in flight, consecutive chunks on the context's two
 * streams; they retire in launch order).  `budget_bytes` of device memory is
 * allocated once; per-request descriptors and output images come from it.
 * submit returns YTRC_EBUDGET when the request does not fit now (no free
 * slot or budget): the caller retries after reading finished requests; and
 * YTRC_EINVAL when it can never fit (more bytes than the whole budget, or a
 * transfer-function table larger than a workgroup's LDS). */
typedef struct ytrc_server ytrc_server;
int ytrc_server_create(ytrc_ctx *ctx, int32_t max_requests, uint64_t budget_bytes,
                       int32_t chunk_tiles, ytrc_server **out);
int ytrc_server_submit(ytrc_server *s, const ytrc_request *r, int64_t *id_out);
/* Retires finished chunks, launches new ones, and returns up to `cap` ids
 * of requests finished since the last call (non-blocking). */
int ytrc_server_progress(ytrc_server *s, int64_t *done_ids, int32_t cap, int32_t *n_done);
/* As progress, but first waits for the oldest chunk in flight. */
int ytrc_server_wait(ytrc_server *s, int64_t *done_ids, int32_t cap, int32_t *n_done);
/* ytrc_server_wait with a deadline: YTRC_ETIMEDOUT (n_done = 0) when no
 * request finished within timeout_ms; YTRC_OK with n_done = 0 when nothing
 * is queued or running. */
int ytrc_server_wait_for(ytrc_server *s, double timeout_ms, int64_t *done_ids, int32_t cap,
                         int32_t *n_done);
/* Copies a finished request's image (float64 nv0*nv1*4) and releases it. */
int ytrc_server_read(ytrc_server *s, int64_t id, double *image_out);
/* Same, as the float32 planes the kernels write (channel-major,
 * nch * nv0 * nv1): the written channels only, no layout conversion. */
int ytrc_server_read_planar(ytrc_server *s, int64_t id, float *planes, int32_t cap_ch, int32_t *nch);
int ytrc_server_cancel(ytrc_server *s, int64_t id);
/* The order in which a server fills its chunks (no counterpart in the
 * reference, which has no queue).  YTRC_SERVE_FIFO (the default): the
 * tiles of admitted requests in arrival order.  YTRC_SERVE_SHORTEST_FIRST:
 * the request with the least estimated work left first (tiles not yet
 * launched; a tile of an interpolating sampler -- volume render,
 * interpolated projection, light source -- counts as 4 projection tiles),
 * except that requests admitted more than `age_limit_ms` ago go before
 * all others, in arrival order, so that none waits without bound.  Takes
 * effect from the next chunk; tiles already launched are not reordered. */
#define YTRC_SERVE_FIFO 0
#define YTRC_SERVE_SHORTEST_FIRST 1
int ytrc_server_set_policy(ytrc_server *s, int32_t policy, double age_limit_ms);
/* stats[8]: requests held, requests with tiles queued, chunks in flight,
 * bytes used, budget, tiles launched, chunks launched, free slots. */
int ytrc_server_stats(const ytrc_server *s, int64_t *stats);
void ytrc_server_destroy(ytrc_server *s);

/* Achievable HBM bandwidth: a streaming copy of `bytes` (read + write
 * counted) timed over `iters` launches with HIP events.  No counterpart in
 * the reference; used to confirm the bandwidth roofline on the device. */
int ytrc_hbm_probe(ytrc_ctx *ctx, uint64_t bytes, int32_t iters, float *gbps, float *ms_per_copy);

/* Average duration (ms) of the dominant kernel over the launches since
 * batch creation, from HIP events on the launch stream. */
int ytrc_batch_kernel_ms(const ytrc_batch *b, float *avg_ms, int32_t *n_launches);

/* kd-tree of an AMR grid hierarchy (host code): the brick set and brick
 * order of the volume render.  Replaces AMRKDTree's build, Tree.build
 * (yt/utilities/amr_kdtree/amr_kdtree.py:63-118) over Node.add_grids /
 * insert_grids / split_grids / divide and kdtree_get_choices
 * (yt/utilities/lib/amr_kdtools.pyx:193-532, :760-823), and its viewpoint
 * traversal kd_traverse(viewpoint) (step_viewpoint, amr_kdtools.pyx:
 * 612-622, 708-755).  Grids are added level by level, lowest first; within
 * a level in the order given (the reference adds data_source.blocks in
 * order). */
typedef struct ytrc_kdtree ytrc_kdtree;
typedef struct {
    double left_edge[3];
    double right_edge[3];
    double split_pos; /* NaN at a leaf */
    int64_t grid;     /* a leaf's grid id; -1 where the node carries none */
    int64_t node_id;  /* 1 at the root, 2 id and 2 id + 1 for the children (mod 2^64) */
    int32_t parent, left, right; /* node indices, -1 where none */
    int32_t split_dim;           /* -1 at a leaf */
} ytrc_kd_node;
int ytrc_kdtree_build(int32_t n_grids, const double *left_edges, const double *right_edges,
                      const int64_t *grid_ids, const int32_t *levels, ytrc_kdtree **out);
/* Node count, and the count of leaves that carry a grid (the bricks). */
int ytrc_kdtree_size(const ytrc_kdtree *t, int32_t *n_nodes, int32_t *n_leaves);
/* All nodes; index 0 is the root, links are indices into this array. */
int ytrc_kdtree_read_nodes(const ytrc_kdtree *t, ytrc_kd_node *nodes);
/* The bricks farthest from `viewpoint` first, as indices into the
 * depth-first (left first) list of leaves that carry a grid: the call order
 * of VolumeSource.render (render_source.py:594-597).  order: n_leaves. */
int ytrc_kdtree_order(const ytrc_kdtree *t, const double *viewpoint, int32_t *order);
void ytrc_kdtree_destroy(ytrc_kdtree *t);

#ifdef __cplusplus
}
#endif
#endif /* YTRC_H */
