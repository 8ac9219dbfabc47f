// ytrc.hip -- libytrc.so: MI355X (gfx950) brick ray-caster behind the C ABI
// declared in include/ytrc.h.
//
// Kernels: one 64-lane wavefront casts one 8x8 tile of rays through the
// bricks of a request in the reference's call order; per brick it applies
// the reference's pixel-extent culling, the FP64 DDA (grid_traversal.pyx
// walk_volume, restated bit-exactly in ytrc_device.h) and the sampler:
//   ProjectionSampler.sample    image_samplers.pyx:330-343
//   VolumeRenderSampler.sample  image_samplers.pyx:453-487
//     + offset_interpolate      fixed_interpolator.cpp:34-52
//     + FIT_eval_transfer       field_interpolation_tables.pxd:66-110
// Data: float32 bricks, fields interleaved per point, packed in one pool.
// Accumulation: projection in FP32 with Kahan compensation; VR compositing
// state in FP64 (full-rate VALU on CDNA4), samples and table lookups FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/ytrc.h"
#include "ytrc_device.h"
#include "ytrc_host.h"
#include "ytrc_kernels.h"
#include "ytrc_pick.h"

// Device memory of ytrc_render_noalloc: every staging buffer of one request
// (its descriptor, brick order, extents, z-buffer, input image, output
// planes, transfer-function table, tile list and queue counters).  Each
// device holds its own copy until the process ends.  In host code the name
// `scratch` is not a device address: host code uses only the address that
// hipGetSymbolAddress returns on each device (ScratchSlot::ptr).
static __device__ __align__(256) char scratch[YTRC_NOALLOC_SCRATCH_BYTES];

using namespace ytrc;

// ---------------------------------------------------------------------------
// Utility kernels

namespace {

// Streaming copy used to measure the achievable HBM bandwidth and to
// calibrate the FETCH_SIZE / WRITE_SIZE counters against a known byte count.
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src,
                                              float4 *__restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Device brick preparation: the vertex data of one kd leaf, computed from its
// grid's cell data (AMRKDTree.get_brick_data, amr_kdtree.py:325-374, with
// Grid.get_vertex_centered_data(no_ghost=True), grid_patch.py:270-296).
struct VertexJob {
    int gdims[3];  // the grid's cells per axis
    int li[3];     // the leaf's first cell in the grid
    int vd[3];     // the leaf's vertices per axis (dims + 1)
    int nb[3];     // the leaf's vertex blocks per axis
    int nf, nfp, logmask;
    int64_t n;     // stored points of the leaf (whole vertex blocks)
    float *out;    // the leaf's points in the pool (nfp floats per point)
};

// Vertex o along an axis of n cells, as ghost_zone_interpolate
// (interpolators.pyx:177-226) places it with rf = 1, input_left = 0.5,
// output_left = 0: cells ii and ii + 1 with weights xm and xp (linear
// extrapolation on the faces).
__device__ __forceinline__ void vc_axis(int o, int n, int &ii, double &xp, double &xm)
{
    const double ropos = ((double)o - 0.5) * 1.0;
    int i = (int)ropos;  // truncation toward zero
    i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
    ii = i;
    xp = ropos - (double)i;
    xm = 1.0 - xp;
}

// One thread per stored point of the leaf (vertex blocks, so a vertex on a
// block face is computed once per block that stores it): the 8-cell
// weighted sum in FP64 in the reference's term order, then log10 with
// v <= 0 -> NaN where the field is a log field (amr_kdtree.py:346-352),
// stored as float32.  Padding points and points past the leaf's last
// vertex are stored as 0 (never read).
template <class T>
__global__ __launch_bounds__(256) void k_vertex_centre(VertexJob J, const T *__restrict__ cells, int64_t nc)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= J.n) return;
    float *o = J.out + p * J.nfp;
    const int local = (int)(p & (VBP - 1));
    const int64_t blk = p >> VBP_SHIFT;
    const int bk = (int)(blk % J.nb[2]);
    const int64_t bq = blk / J.nb[2];
    const int bj = (int)(bq % J.nb[1]);
    const int bi = (int)(bq / J.nb[1]);
    const int a = bi * VB + local / (VB1 * VB1);
    const int b = bj * VB + (local / VB1) % VB1;
    const int c = bk * VB + local % VB1;
    if (local >= VB1 * VB1 * VB1 || a >= J.vd[0] || b >= J.vd[1] || c >= J.vd[2]) {
        for (int f = 0; f < J.nfp; f++) o[f] = 0.0f;
        return;
    }
    int ii, ij, ik;
    double xp, xm, yp, ym, zp, zm;
    vc_axis(J.li[0] + a, J.gdims[0], ii, xp, xm);
    vc_axis(J.li[1] + b, J.gdims[1], ij, yp, ym);
    vc_axis(J.li[2] + c, J.gdims[2], ik, zp, zm);
    const int64_t s1 = J.gdims[2];
    const int64_t s0 = (int64_t)J.gdims[1] * J.gdims[2];
    for (int f = 0; f < J.nfp; f++) {
        if (f >= J.nf) {
            o[f] = 0.0f;
            continue;
        }
        const T *g = cells + (int64_t)f * nc + ii * s0 + ij * s1 + ik;
        double v = (double)g[0] * (xm * ym * zm);
        v = v + (double)g[s0] * (xp * ym * zm);
        v = v + (double)g[s1] * (xm * yp * zm);
        v = v + (double)g[1] * (xm * ym * zp);
        v = v + (double)g[s0 + 1] * (xp * ym * zp);
        v = v + (double)g[s1 + 1] * (xm * yp * zp);
        v = v + (double)g[s0 + s1] * (xp * yp * zm);
        v = v + (double)g[s0 + s1 + 1] * (xp * yp * zp);
        if ((J.logmask >> f) & 1) v = (v <= 0.0) ? __builtin_nan("") : log10(v);
        o[f] = (float)v;
    }
}

__global__ void k_zero_counters(unsigned int *c, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0u;
}

// Stores field f of one brick, uploaded in C order (src, or none for the
// padding fields f >= nf), in the device layout: one thread per stored
// point; 4x4x4 cell blocks for cell-centred bricks, vertex blocks for
// vertex-centred ones (a face vertex is read once per block that stores
// it).  Points of no cell or vertex (block padding) are stored as 0.
template <class T>
__global__ __launch_bounds__(256) void k_store_field(const T *__restrict__ src, float *__restrict__ dst,
                                                      int nfp, int f, int d0, int d1, int d2,
                                                      int vertex, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int64_t s = -1;
    if (vertex) {
        const int local = (int)(p & (VBP - 1));
        const int64_t blk = p >> VBP_SHIFT;
        const int nb1 = vblocks(d1), nb2 = vblocks(d2);
        const int bk = (int)(blk % nb2);
        const int64_t bq = blk / nb2;
        const int bj = (int)(bq % nb1);
        const int bi = (int)(bq / nb1);
        const int a = bi * VB + local / (VB1 * VB1);
        const int b = bj * VB + (local / VB1) % VB1;
        const int c = bk * VB + local % VB1;
        if (local < VB1 * VB1 * VB1 && a <= d0 && b <= d1 && c <= d2)
            s = ((int64_t)a * (d1 + 1) + b) * (d2 + 1) + c;
    } else {
        const int local = (int)(p & 63);
        const int64_t blk = p >> 6;
        const int nbj = (d1 + 3) >> 2, nbk = (d2 + 3) >> 2;
        const int bk = (int)(blk % nbk);
        const int64_t bq = blk / nbk;
        const int bj = (int)(bq % nbj);
        const int bi = (int)(bq / nbj);
        const int i = 4 * bi + (local >> 4), j = 4 * bj + ((local >> 2) & 3), k = 4 * bk + (local & 3);
        if (i < d0 && j < d1 && k < d2) s = ((int64_t)i * d1 + j) * d2 + k;
    }
    dst[p * nfp + f] = (src != nullptr && s >= 0) ? (float)src[s] : 0.0f;
}

}  // namespace

namespace {

// Kernel variant of a request: KIND, interleaved fields NFP, record mode and
// MODE (VR: bit 0 ColorTransferFunction path, bit 1 n_samples == 10, bit 2
// early ray termination, bit 3 per-pixel lens; projection: bit 3).  NFP 4
// and 8 variants keep only bits 0 and 3.
// Sampler kinds that read a transfer function.
inline bool has_tf(int kind) { return kind == YTRC_VOLUME_RENDER || kind == YTRC_LIGHT_SOURCE; }

int variant_mode(int kind, int nfp, int mode)
{
    if (nfp >= 4) return (mode & 4) ? -1 : (mode & 9);
    (void)kind;
    return mode;
}

// The tiles of an image of tiles0 x tiles1 ray tiles in Z order (the bits
// of the tile row and column interleaved), so that every run of the
// sequence covers a compact patch of the image.
std::vector<int> tiles_z_order(int tiles0, int tiles1)
{
    auto spread = [](uint32_t v) {
        uint64_t x = v;
        x = (x | (x << 16)) & 0x0000ffff0000ffffull;
        x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
        x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
        x = (x | (x << 2)) & 0x3333333333333333ull;
        x = (x | (x << 1)) & 0x5555555555555555ull;
        return x;
    };
    std::vector<std::pair<uint64_t, int>> keys((size_t)tiles0 * tiles1);
    for (int a = 0; a < tiles0; a++)
        for (int b = 0; b < tiles1; b++)
            keys[(size_t)a * tiles1 + b] = {(spread((uint32_t)a) << 1) | spread((uint32_t)b), a * tiles1 + b};
    std::sort(keys.begin(), keys.end());
    std::vector<int> order(keys.size());
    for (size_t k = 0; k < keys.size(); k++) order[k] = keys[k].second;
    return order;
}

// Words of LDS per wave that a persistent launch gives after its tables
// (k_persist's wave_words): the vertex-block buffer with mode bit 4, a
// projection's table of offset terms (walk_projection_fast).
int required_wave_words(int kind, int nfp, bool rec, int mode)
{
    if (mode & 16) return VBP;
    if (kind == YTRC_PROJECTION && !rec && nfp <= 4) return PTAB_WAVE;
    return 0;
}

pfn_t pick_persist(int kind, int nfp, bool rec, int mode)
{
    const int m = variant_mode(kind, nfp, mode);
    if (m < 0) return nullptr;
    if (kind == 0) return pick_persist_proj(nfp, rec, m);
    if (kind == 2) return pick_persist_iproj(nfp, rec, m);
    if (kind == 3) return pick_persist_light(nfp, rec, m);
    if (m < 2) return pick_persist_vr_a(nfp, rec, m);
    if (m < 4 || m == 19) return pick_persist_vr_b(nfp, rec, m);
    if (m < 8) return pick_persist_vr_c(nfp, rec, m);
    return pick_persist_vr_d(nfp, rec, m);
}

kfn_t pick_kernel(int kind, int nfp, bool rec, int mode)
{
    const int m = variant_mode(kind, nfp, mode);
    if (m < 0) return nullptr;
    if (kind == 0) return pick_kernel_proj(nfp, rec, m);
    if (kind == 2) return pick_kernel_iproj(nfp, rec, m);
    if (kind == 3) return pick_kernel_light(nfp, rec, m);
    if (m < 2) return pick_kernel_vr_a(nfp, rec, m);
    if (m < 4) return pick_kernel_vr_b(nfp, rec, m);
    if (m < 8) return pick_kernel_vr_c(nfp, rec, m);
    return pick_kernel_vr_d(nfp, rec, m);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

}  // namespace

int ytrc::set_error(int code, const std::string &msg) { return fail(code, msg); }

namespace {

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(YTRC_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));   \
    } while (0)

int nfp_of(int nf) { return nf <= 1 ? 1 : (nf == 2 ? 2 : (nf <= 4 ? 4 : 8)); }

struct Dataset {
    int n_bricks = 0;
    int nf = 0;
    int nfp = 0;
    std::vector<DBrick> h_bricks;
    DBrick *d_bricks = nullptr;
    float *d_pool = nullptr;
    uint8_t *d_masks = nullptr;
    uint64_t bytes = 0;
    int refs = 0;  // live batches and server requests that read it
};

}  // namespace

struct ytrc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // a second stream (default priority) for the concurrent launches of a
    // persistent batch, joined back to `stream` by the batch's events
    hipStream_t stream2 = nullptr;
    uint64_t budget = 0;
    uint64_t used = 0;
    int n_cu = 0;
    size_t max_lds = 64 * 1024;  // LDS bytes one workgroup may allocate
    std::string name;            // marketing name and ISA of the device
    char *scratch = nullptr;     // this device's `scratch` array (ytrc_render_noalloc)
    std::mutex *scratch_mu = nullptr;  // serialises the calls that share it
    std::map<int, Dataset> datasets;
    int next_ds = 1;
    std::mutex mu;
};

namespace {

// Projection samplers add to the image (image_samplers.pyx:330-343,
// 368-402): the device integrates from zero and the caller's float64
// image_in is added on the host at read-back, so a pixel no visit touches
// keeps its input bits.  Compositing samplers start from image_in on the
// device.
static bool additive_kind(int32_t kind)
{
    return kind == YTRC_PROJECTION || kind == YTRC_INTERPOLATED_PROJECTION;
}

struct ReqHost {
    ytrc_request r;
    std::vector<double> image_in;  // nv0*nv1*4 or empty
    int64_t npix = 0;
    int ntiles = 0;
    kfn_t kfn = nullptr;
    size_t lds = 0;
    int nch_out = 0;
    int nfp = 1;
    int mode = 0;
    int tform = 0;
};

}  // namespace

// One persistent launch: every request of one kernel variant.
struct PGroup {
    int kind = 0;
    int mode = 0;
    int nfp = 1;
    bool rec = false;
    pfn_t fn = nullptr;
    int n_items = 0;
    const int2 *items = nullptr;
    int counter = 0;
    size_t lds = 0;
    int grid = 0;
    int tf_base = 0;  // the launch's tables: float4 offset in the batch table
    int tf_bins = 0;  // and length (what each workgroup stages in LDS)
};

struct ytrc_batch {
    ytrc_ctx *ctx = nullptr;
    uint32_t flags = 0;
    std::vector<ReqHost> reqs;
    std::vector<DImg> h_imgs;
    DImg *d_imgs = nullptr;
    char *arena = nullptr;
    uint64_t arena_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;  // persistent: the second stream's fork and join
    std::vector<hipEvent_t> ev_req;  // plain mode: before/after each request
    std::vector<PGroup> groups;      // persistent mode
    unsigned int *counters = nullptr;
    const float4 *tf_all = nullptr;
    int tf_bins = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int waves_per_group = 4;
    int groups_per_cu = 0;  // 0: occupancy limit
    int launched = 0;
    double total_ms = 0.0;
    int timed = 0;
    std::vector<int> ds_ids;  // datasets this batch holds a reference on
    int *h_stop = nullptr;    // host-pinned stop flag read by the kernels (ytrc_batch_cancel)
    bool cancelled = false;   // the last launch was asked to stop: its images are incomplete
};

namespace {

// calculate_extent_plane_parallel (lenses.pyx:22-55) + the +-1 clip of
// cast_through_kdtree (image_samplers.pyx:159-166), on the host, in FP64.
void extent_plane_parallel(const ytrc_request &r, const DBrick &b, int64_t rv[4])
{
    const double *edges[2] = {b.le, b.re};
    double ext[4] = {1e300, -1e300, 1e300, -1e300};
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) {
                double t = edges[i][0] * r.x_vec[0];
                t += edges[j][1] * r.x_vec[1];
                t += edges[k][2] * r.x_vec[2];
                if (t < ext[0]) ext[0] = t;
                if (t > ext[1]) ext[1] = t;
                t = edges[i][0] * r.y_vec[0];
                t += edges[j][1] * r.y_vec[1];
                t += edges[k][2] * r.y_vec[2];
                if (t < ext[2]) ext[2] = t;
                if (t > ext[3]) ext[3] = t;
            }
    double cx = 0.0, cy = 0.0;
    for (int i = 0; i < 3; i++) {
        cx += r.center[i] * r.x_vec[i];
        cy += r.center[i] * r.y_vec[i];
    }
    const double pdx = (r.bounds[1] - r.bounds[0]) / (double)r.nv[0];
    const double pdy = (r.bounds[3] - r.bounds[2]) / (double)r.nv[1];
    rv[0] = lrint((ext[0] - cx - r.bounds[0]) / pdx);
    rv[1] = rv[0] + lrint((ext[1] - ext[0]) / pdx);
    rv[2] = lrint((ext[2] - cy - r.bounds[2]) / pdy);
    rv[3] = rv[2] + lrint((ext[3] - ext[2]) / pdy);
    auto clip = [](int64_t v, int64_t a, int64_t bnd) { return v < a ? a : (v > bnd ? bnd : v); };
    rv[0] = clip(rv[0] - 1, 0, r.nv[0]);
    rv[1] = clip(rv[1] + 1, 0, r.nv[0]);
    rv[2] = clip(rv[2] - 1, 0, r.nv[1]);
    rv[3] = clip(rv[3] + 1, 0, r.nv[1]);
}

// calculate_extent_perspective (lenses.pyx:60-161) + the +-1 clip, on the
// host.  The reference declares sight_vector_norm, sight_angle_cos,
// sight_length, dx and dy as C float: each is rounded to float32 where it
// is assigned; everything else is double (vec3_ops.pxd's double versions).
void extent_perspective(const ytrc_request &r, const DBrick &b, int64_t rv[4])
{
    const int nv0 = r.nv[0], nv1 = r.nv[1];
    const double *cd = r.camera_data;
    double cam_pos[3], cam_width[3], east[3], north[3], normal[3];
    for (int i = 0; i < 3; i++) {
        cam_pos[i] = cd[i];
        cam_width[i] = cd[3 + i];
        east[i] = cd[6 + i];
        north[i] = cd[9 + i];
        normal[i] = cd[12 + i];
    }
    // corner order of lenses.pyx:80-103
    const int cx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
    const int cy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    const int cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    int min_px = 32767, min_py = 32767, max_px = -32767, max_py = -32767;
    for (int iv = 0; iv < 8; iv++) {
        const double vertex[3] = {cx[iv] ? b.re[0] : b.le[0], cy[iv] ? b.re[1] : b.le[1],
                                  cz[iv] ? b.re[2] : b.le[2]};
        cam_width[1] = cam_width[0] * nv1 / nv0;
        double sv[3], sc[3], pos1[3];
        for (int i = 0; i < 3; i++) sv[i] = vertex[i] - cam_pos[i];
        for (int i = 0; i < 3; i++) sc[i] = cam_width[2] * normal[i] + cam_pos[i];
        const float norm = (float)std::sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
        if (norm != 0)
            for (int i = 0; i < 3; i++) sv[i] /= (double)norm;
        float cosa = (float)(sv[0] * normal[0] + sv[1] * normal[1] + sv[2] * normal[2]);
        cosa = (float)std::fmin(std::fmax((double)cosa, -1.0), 1.0);
        float len;
        if (std::acos((double)cosa) < 0.5 * 3.14159265358979323846 && cosa != 0.0) {
            len = (float)(cam_width[2] / (double)cosa);
        } else {
            len = (float)std::sqrt(cam_width[0] * cam_width[0] + cam_width[1] * cam_width[1]);
            const float c2 = cosa * cosa;
            len = (float)((double)len / std::sqrt(1.0 - (double)c2));
        }
        for (int i = 0; i < 3; i++) pos1[i] = (double)len * sv[i] + cam_pos[i];
        for (int i = 0; i < 3; i++) pos1[i] = pos1[i] - sc[i];
        const float dx = (float)(pos1[0] * east[0] + pos1[1] * east[1] + pos1[2] * east[2]);
        const float dy = (float)(pos1[0] * north[0] + pos1[1] * north[1] + pos1[2] * north[2]);
        const int px = (int)(nv0 * 0.5 + nv0 / cam_width[0] * (double)dx);
        const int py = (int)(nv1 * 0.5 + nv1 / cam_width[1] * (double)dy);
        min_px = std::min(min_px, px);
        max_px = std::max(max_px, px);
        min_py = std::min(min_py, py);
        max_py = std::max(max_py, py);
    }
    rv[0] = std::max(min_px, 0);
    rv[1] = std::min(max_px, nv0);
    rv[2] = std::max(min_py, 0);
    rv[3] = std::min(max_py, nv1);
    auto clip = [](int64_t v, int64_t a, int64_t bnd) { return v < a ? a : (v > bnd ? bnd : v); };
    rv[0] = clip(rv[0] - 1, 0, nv0);
    rv[1] = clip(rv[1] + 1, 0, nv0);
    rv[2] = clip(rv[2] - 1, 0, nv1);
    rv[3] = clip(rv[3] + 1, 0, nv1);
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

namespace {

void dataset_release(ytrc_ctx *ctx, Dataset &ds);

// Validates brick descriptors, lays out the point pool (stored_points per
// brick) and the masks, allocates device memory, and copies descriptors and
// masks.  Field data is filled in by the caller.
int dataset_alloc(ytrc_ctx *ctx, const char *who, int n_bricks, const ytrc_brick_desc *bricks,
                  int n_fields, const uint8_t *const *mask_ptrs, Dataset &ds)
{
    const std::string w(who);
    if (n_bricks <= 0) return fail(YTRC_EINVAL, w + ": n_bricks must be > 0");
    if (n_fields <= 0 || n_fields > 6) return fail(YTRC_EINVAL, w + ": n_fields must be in 1..6");
    ds.n_bricks = n_bricks;
    ds.nf = n_fields;
    ds.nfp = nfp_of(n_fields);
    ds.h_bricks.resize(n_bricks);
    int64_t pool_floats = 0, mask_bytes = 0;
    for (int b = 0; b < n_bricks; b++) {
        const ytrc_brick_desc &d = bricks[b];
        DBrick &h = ds.h_bricks[b];
        memset(&h, 0, sizeof(h));
        for (int i = 0; i < 3; i++) {
            if (d.dims[i] <= 0)
                return fail(YTRC_EINVAL, w + ": brick " + std::to_string(b) + " has non-positive dims");
            if (!(d.right_edge[i] > d.left_edge[i]))
                return fail(YTRC_EINVAL, w + ": brick " + std::to_string(b) + " has right_edge <= left_edge");
            h.le[i] = d.left_edge[i];
            h.re[i] = d.right_edge[i];
            h.dims[i] = d.dims[i];
            // partitioned_grid.pyx:50-55
            h.dds[i] = (h.re[i] - h.le[i]) / (double)d.dims[i];
            h.idds[i] = 1.0 / h.dds[i];
        }
        h.vertex = d.vertex_centred ? 1 : 0;
        if (stored_points(h) >= (int64_t)1 << 31)
            return fail(YTRC_EINVAL, w + ": brick " + std::to_string(b) +
                                         " exceeds 2^31 points (int32 cell index, image_samplers.pyx:341)");
        // The kernels form cell and vertex offsets with 24-bit multiplies.
        if ((int64_t)(d.dims[0] + 1) * (d.dims[1] + 1) >= (int64_t)1 << 24 ||
            (int64_t)(d.dims[1] + 1) * (d.dims[2] + 1) >= (int64_t)1 << 24)
            return fail(YTRC_EINVAL, w + ": brick " + std::to_string(b) +
                                         " has a face of 2^24 or more points; split it into smaller bricks");
        h.data_off = pool_floats;
        pool_floats += align_up((size_t)stored_points(h) * ds.nfp, 64);
        bool has_mask = mask_ptrs != nullptr && mask_ptrs[b] != nullptr;
        if (has_mask) {
            // An all-ones mask is the same as none: skip its per-visit load.
            const size_t nc = (size_t)d.dims[0] * d.dims[1] * d.dims[2];
            const uint8_t *m = mask_ptrs[b];
            bool all1 = true;
            for (size_t q = 0; q < nc && all1; q++) all1 = (m[q] == 1);
            has_mask = !all1;
        }
        h.mask_off = has_mask ? mask_bytes : -1;
        if (has_mask) mask_bytes += align_up((size_t)d.dims[0] * d.dims[1] * d.dims[2], 256);
    }
    const uint64_t bytes = (uint64_t)pool_floats * 4 + (uint64_t)mask_bytes +
                           (uint64_t)n_bricks * sizeof(DBrick);
    if (ctx->budget != 0 && ctx->used + bytes > ctx->budget)
        return fail(YTRC_EBUDGET, w + ": " + std::to_string(bytes) + " bytes exceed the HBM budget");
    auto dev = [&](hipError_t e, const char *what) {
        dataset_release(ctx, ds);
        return fail(YTRC_EDEVICE, w + ": " + what + ": " + hipGetErrorString(e));
    };
    hipError_t e = hipMalloc(&ds.d_pool, (size_t)pool_floats * 4);
    if (e != hipSuccess) return dev(e, "hipMalloc(pool)");
    if (mask_bytes > 0) {
        e = hipMalloc(&ds.d_masks, (size_t)mask_bytes);
        if (e != hipSuccess) return dev(e, "hipMalloc(masks)");
    }
    e = hipMalloc(&ds.d_bricks, (size_t)n_bricks * sizeof(DBrick));
    if (e != hipSuccess) return dev(e, "hipMalloc(bricks)");
    for (int b = 0; b < n_bricks; b++) {
        const DBrick &h = ds.h_bricks[b];
        if (h.mask_off < 0) continue;
        e = hipMemcpy(ds.d_masks + h.mask_off, mask_ptrs[b], (size_t)h.dims[0] * h.dims[1] * h.dims[2],
                      hipMemcpyHostToDevice);
        if (e != hipSuccess) return dev(e, "hipMemcpy(mask)");
    }
    e = hipMemcpy(ds.d_bricks, ds.h_bricks.data(), (size_t)n_bricks * sizeof(DBrick), hipMemcpyHostToDevice);
    if (e != hipSuccess) return dev(e, "hipMemcpy(bricks)");
    ds.bytes = bytes;
    return YTRC_OK;
}

void dataset_release(ytrc_ctx *ctx, Dataset &ds)
{
    (void)ctx;
    if (ds.d_bricks) hipFree(ds.d_bricks);
    if (ds.d_pool) hipFree(ds.d_pool);
    if (ds.d_masks) hipFree(ds.d_masks);
    ds.d_bricks = nullptr;
    ds.d_pool = nullptr;
    ds.d_masks = nullptr;
}

int dataset_commit(ytrc_ctx *ctx, Dataset &ds, int32_t *ds_id_out)
{
    ctx->used += ds.bytes;
    const int id = ctx->next_ds++;
    ctx->datasets[id] = std::move(ds);
    *ds_id_out = id;
    return YTRC_OK;
}

}  // namespace

namespace {

// The address of `scratch` on each device, looked up once per device with
// that device current, and a mutex per device for the calls sharing it.
struct ScratchSlot {
    char *ptr = nullptr;
    std::mutex mu;
};
std::mutex g_scratch_map_mu;
std::map<int, ScratchSlot> g_scratch;

int scratch_setup(ytrc_ctx *c)
{
    std::lock_guard<std::mutex> lk(g_scratch_map_mu);
    ScratchSlot &slot = g_scratch[c->device];
    if (slot.ptr == nullptr) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr)
            return fail(YTRC_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        slot.ptr = (char *)ptr;
    }
    c->scratch = slot.ptr;
    c->scratch_mu = &slot.mu;
    return YTRC_OK;
}

}  // namespace

extern "C" {

const char *ytrc_last_error(void) { return g_err.c_str(); }

int ytrc_ctx_create(int32_t device, uint64_t budget_bytes, ytrc_ctx **out)
{
    if (out == nullptr) return fail(YTRC_EINVAL, "ytrc_ctx_create: out is NULL");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(YTRC_EINVAL, "ytrc_ctx_create: no such device " + std::to_string(device));
    HIPCHK(hipSetDevice(device));
    ytrc_ctx *c = new ytrc_ctx();
    c->device = device;
    c->budget = budget_bytes;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete c;
        return fail(YTRC_EDEVICE, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    }
    c->n_cu = prop.multiProcessorCount;
    if (prop.sharedMemPerBlock > 0) c->max_lds = prop.sharedMemPerBlock;
    c->name = std::string(prop.name) + " (" + prop.gcnArchName + ", " +
              std::to_string(prop.multiProcessorCount) + " CUs)";
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(YTRC_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    if ((e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking)) != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        return fail(YTRC_EDEVICE, std::string("ytrc_ctx_create: second stream: ") + hipGetErrorString(e));
    }
    if (scratch_setup(c) != YTRC_OK) {
        hipStreamDestroy(c->stream);
        delete c;
        return YTRC_EDEVICE;
    }
    *out = c;
    return YTRC_OK;
}

void ytrc_ctx_destroy(ytrc_ctx *ctx)
{
    if (ctx == nullptr) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    hipStreamSynchronize(ctx->stream2);
    for (auto &kv : ctx->datasets) {
        hipFree(kv.second.d_bricks);
        hipFree(kv.second.d_pool);
        if (kv.second.d_masks) hipFree(kv.second.d_masks);
    }
    hipStreamDestroy(ctx->stream2);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

int ytrc_ctx_max_lds(const ytrc_ctx *ctx, uint64_t *bytes)
{
    if (ctx == nullptr || bytes == nullptr) return fail(YTRC_EINVAL, "ytrc_ctx_max_lds: NULL");
    *bytes = ctx->max_lds;
    return YTRC_OK;
}

int ytrc_ctx_device_name(const ytrc_ctx *ctx, char *buf, int32_t cap)
{
    if (ctx == nullptr || buf == nullptr || cap <= 0)
        return fail(YTRC_EINVAL, "ytrc_ctx_device_name: NULL or empty buffer");
    const size_t n = std::min(ctx->name.size(), (size_t)cap - 1);
    memcpy(buf, ctx->name.data(), n);
    buf[n] = '\0';
    return YTRC_OK;
}

int ytrc_ctx_used_bytes(const ytrc_ctx *ctx, uint64_t *used, uint64_t *budget)
{
    if (ctx == nullptr) return fail(YTRC_EINVAL, "ytrc_ctx_used_bytes: ctx is NULL");
    if (used) *used = ctx->used;
    if (budget) *budget = ctx->budget;
    return YTRC_OK;
}

int ytrc_dataset_upload(ytrc_ctx *ctx, int32_t n_bricks, const ytrc_brick_desc *bricks,
                        int32_t n_fields, int32_t src_dtype, const void *const *field_ptrs,
                        const uint8_t *const *mask_ptrs, int32_t *ds_id_out)
{
    if (ctx == nullptr || bricks == nullptr || field_ptrs == nullptr || ds_id_out == nullptr)
        return fail(YTRC_EINVAL, "ytrc_dataset_upload: NULL argument");
    if (src_dtype != 0 && src_dtype != 1)
        return fail(YTRC_EINVAL, "ytrc_dataset_upload: src_dtype must be 0 (f64) or 1 (f32)");
    for (int64_t q = 0; q < (int64_t)n_bricks * std::max(0, n_fields); q++)
        if (field_ptrs[q] == nullptr) return fail(YTRC_EINVAL, "ytrc_dataset_upload: NULL field pointer");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    Dataset ds;
    int rc = dataset_alloc(ctx, "ytrc_dataset_upload", n_bricks, bricks, n_fields, mask_ptrs, ds);
    if (rc != YTRC_OK) return rc;
    // Each brick field goes to the device in C order as given (float64 or
    // float32) and k_store_field converts it to float32 in the device
    // layout, fields interleaved per point.  One staging buffer holds the
    // largest brick field; copies and kernels are ordered on the stream.
    const size_t esz = src_dtype == 0 ? 8 : 4;
    int64_t max_np = 0;
    for (int b = 0; b < n_bricks; b++) {
        const DBrick &h = ds.h_bricks[b];
        max_np = std::max(max_np, (int64_t)(h.dims[0] + h.vertex) * (h.dims[1] + h.vertex) *
                                      (h.dims[2] + h.vertex));
    }
    void *d_src = nullptr;
    hipError_t e = hipMalloc(&d_src, (size_t)max_np * esz);
    if (e != hipSuccess) {
        dataset_release(ctx, ds);
        return fail(YTRC_EDEVICE, std::string("hipMalloc(staging): ") + hipGetErrorString(e));
    }
    for (int b = 0; b < n_bricks && e == hipSuccess; b++) {
        const DBrick &h = ds.h_bricks[b];
        const int64_t np = (int64_t)(h.dims[0] + h.vertex) * (h.dims[1] + h.vertex) *
                           (h.dims[2] + h.vertex);
        const int64_t n = stored_points(h);
        const unsigned grid = (unsigned)((n + 255) / 256);
        for (int f = 0; f < ds.nfp && e == hipSuccess; f++) {
            const bool have = f < n_fields;
            if (have)
                e = hipMemcpyAsync(d_src, field_ptrs[(size_t)b * n_fields + f], (size_t)np * esz,
                                   hipMemcpyHostToDevice, ctx->stream);
            if (e != hipSuccess) break;
            if (src_dtype == 0)
                hipLaunchKernelGGL(k_store_field<double>, dim3(grid), dim3(256), 0, ctx->stream,
                                   have ? (const double *)d_src : nullptr, ds.d_pool + h.data_off,
                                   ds.nfp, f, h.dims[0], h.dims[1], h.dims[2], h.vertex, n);
            else
                hipLaunchKernelGGL(k_store_field<float>, dim3(grid), dim3(256), 0, ctx->stream,
                                   have ? (const float *)d_src : nullptr, ds.d_pool + h.data_off,
                                   ds.nfp, f, h.dims[0], h.dims[1], h.dims[2], h.vertex, n);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d_src);
    if (e != hipSuccess) {
        dataset_release(ctx, ds);
        return fail(YTRC_EDEVICE, std::string("ytrc_dataset_upload: ") + hipGetErrorString(e));
    }
    return dataset_commit(ctx, ds, ds_id_out);
}

int ytrc_dataset_build_vr(ytrc_ctx *ctx, int32_t n_grids, const ytrc_grid_desc *grids,
                          int32_t n_fields, int32_t src_dtype, const void *const *cell_ptrs,
                          const int32_t *log_fields, int32_t n_leaves, const ytrc_leaf_desc *leaves,
                          const uint8_t *const *mask_ptrs, int32_t *ds_id_out)
{
    const char *who = "ytrc_dataset_build_vr";
    if (ctx == nullptr || grids == nullptr || cell_ptrs == nullptr || leaves == nullptr ||
        ds_id_out == nullptr)
        return fail(YTRC_EINVAL, std::string(who) + ": NULL argument");
    if (n_grids <= 0) return fail(YTRC_EINVAL, std::string(who) + ": n_grids must be > 0");
    if (src_dtype != 0 && src_dtype != 1)
        return fail(YTRC_EINVAL, std::string(who) + ": src_dtype must be 0 (f64) or 1 (f32)");
    if (n_fields <= 0 || n_fields > 6)
        return fail(YTRC_EINVAL, std::string(who) + ": n_fields must be in 1..6");
    // Grid cell data, back to back per (grid, field), in the source dtype.
    const size_t esz = src_dtype == 0 ? 8 : 4;
    std::vector<int64_t> goff((size_t)n_grids);
    int64_t cells = 0;
    for (int g = 0; g < n_grids; g++) {
        for (int a = 0; a < 3; a++)
            if (grids[g].dims[a] < 2)
                return fail(YTRC_EINVAL, std::string(who) + ": grid " + std::to_string(g) +
                                             " needs at least 2 cells per axis (interpolators.pyx:177-226)");
        goff[(size_t)g] = cells;
        cells += (int64_t)grids[g].dims[0] * grids[g].dims[1] * grids[g].dims[2] * n_fields;
    }
    for (int64_t q = 0; q < (int64_t)n_grids * n_fields; q++)
        if (cell_ptrs[q] == nullptr) return fail(YTRC_EINVAL, std::string(who) + ": NULL cell pointer");
    // Leaves become vertex-centred brick descriptors (PartitionedGrid(...,
    // node edges, dims = ri - li), amr_kdtree.py:360-367).
    std::vector<ytrc_brick_desc> bd((size_t)std::max(0, n_leaves));
    for (int l = 0; l < n_leaves; l++) {
        const ytrc_leaf_desc &L = leaves[l];
        if (L.grid < 0 || L.grid >= n_grids)
            return fail(YTRC_EINVAL, std::string(who) + ": leaf " + std::to_string(l) + " names no grid");
        for (int a = 0; a < 3; a++) {
            if (L.li[a] < 0 || L.ri[a] > grids[L.grid].dims[a] || L.ri[a] <= L.li[a])
                return fail(YTRC_EINVAL, std::string(who) + ": leaf " + std::to_string(l) +
                                             " slice lies outside its grid");
            bd[(size_t)l].left_edge[a] = L.left_edge[a];
            bd[(size_t)l].right_edge[a] = L.right_edge[a];
            bd[(size_t)l].dims[a] = L.ri[a] - L.li[a];
        }
        bd[(size_t)l].vertex_centred = 1;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    Dataset ds;
    int rc = dataset_alloc(ctx, who, n_leaves, bd.data(), n_fields, mask_ptrs, ds);
    if (rc != YTRC_OK) return rc;
    void *d_cells = nullptr;
    hipError_t e = hipMalloc(&d_cells, (size_t)cells * esz);
    if (e != hipSuccess) {
        dataset_release(ctx, ds);
        return fail(YTRC_EDEVICE, std::string("hipMalloc(cells): ") + hipGetErrorString(e));
    }
    auto bail = [&](hipError_t err, const char *what) {
        hipFree(d_cells);
        dataset_release(ctx, ds);
        return fail(YTRC_EDEVICE, std::string(what) + ": " + hipGetErrorString(err));
    };
    for (int g = 0; g < n_grids; g++) {
        const size_t nc = (size_t)grids[g].dims[0] * grids[g].dims[1] * grids[g].dims[2];
        for (int f = 0; f < n_fields; f++) {
            e = hipMemcpyAsync((char *)d_cells + ((size_t)goff[(size_t)g] + (size_t)f * nc) * esz,
                               cell_ptrs[(size_t)g * n_fields + f], nc * esz, hipMemcpyHostToDevice,
                               ctx->stream);
            if (e != hipSuccess) return bail(e, "hipMemcpyAsync(cells)");
        }
    }
    int logmask = 0;
    for (int f = 0; f < n_fields; f++)
        if (log_fields != nullptr && log_fields[f]) logmask |= 1 << f;
    for (int l = 0; l < n_leaves; l++) {
        const ytrc_leaf_desc &L = leaves[l];
        const DBrick &h = ds.h_bricks[(size_t)l];
        VertexJob J;
        for (int a = 0; a < 3; a++) {
            J.gdims[a] = grids[L.grid].dims[a];
            J.li[a] = L.li[a];
            J.vd[a] = h.dims[a] + 1;
            J.nb[a] = vblocks(h.dims[a]);
        }
        J.nf = n_fields;
        J.nfp = ds.nfp;
        J.logmask = logmask;
        J.n = stored_points(h);
        J.out = ds.d_pool + h.data_off;
        const int64_t nc = (int64_t)J.gdims[0] * J.gdims[1] * J.gdims[2];
        const unsigned grid = (unsigned)((J.n + 255) / 256);
        if (src_dtype == 0)
            hipLaunchKernelGGL(k_vertex_centre<double>, dim3(grid), dim3(256), 0, ctx->stream, J,
                               (const double *)d_cells + goff[(size_t)L.grid], nc);
        else
            hipLaunchKernelGGL(k_vertex_centre<float>, dim3(grid), dim3(256), 0, ctx->stream, J,
                               (const float *)d_cells + goff[(size_t)L.grid], nc);
        e = hipGetLastError();
        if (e != hipSuccess) return bail(e, "k_vertex_centre");
    }
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return bail(e, "hipStreamSynchronize");
    hipFree(d_cells);
    return dataset_commit(ctx, ds, ds_id_out);
}

int ytrc_dataset_read_points(ytrc_ctx *ctx, int32_t ds_id, int32_t brick, float *out, int64_t cap,
                             int64_t *n_out)
{
    if (ctx == nullptr || n_out == nullptr) return fail(YTRC_EINVAL, "ytrc_dataset_read_points: NULL");
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->datasets.find(ds_id);
    if (it == ctx->datasets.end())
        return fail(YTRC_EINVAL, "ytrc_dataset_read_points: unknown dataset " + std::to_string(ds_id));
    const Dataset &ds = it->second;
    if (brick < 0 || brick >= ds.n_bricks)
        return fail(YTRC_EINVAL, "ytrc_dataset_read_points: brick index out of range");
    const DBrick &h = ds.h_bricks[(size_t)brick];
    const int64_t n = stored_points(h) * ds.nfp;
    *n_out = n;
    if (out == nullptr) return YTRC_OK;
    if (cap < n) return fail(YTRC_EINVAL, "ytrc_dataset_read_points: output too small");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(out, ds.d_pool + h.data_off, (size_t)n * 4, hipMemcpyDeviceToHost));
    return YTRC_OK;
}

int ytrc_dataset_free(ytrc_ctx *ctx, int32_t ds_id)
{
    if (ctx == nullptr) return fail(YTRC_EINVAL, "ytrc_dataset_free: ctx is NULL");
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->datasets.find(ds_id);
    if (it == ctx->datasets.end())
        return fail(YTRC_EINVAL, "ytrc_dataset_free: unknown dataset " + std::to_string(ds_id));
    if (it->second.refs > 0)
        return fail(YTRC_EBUSY, "ytrc_dataset_free: dataset " + std::to_string(ds_id) + " is used by " +
                                    std::to_string(it->second.refs) +
                                    " live batch(es) or server request(s); destroy or read them first");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream2));
    hipFree(it->second.d_bricks);
    hipFree(it->second.d_pool);
    if (it->second.d_masks) hipFree(it->second.d_masks);
    ctx->used -= it->second.bytes;
    ctx->datasets.erase(it);
    return YTRC_OK;
}

int ytrc_dataset_bytes(const ytrc_ctx *ctx, int32_t ds_id, uint64_t *bytes)
{
    if (ctx == nullptr || bytes == nullptr) return fail(YTRC_EINVAL, "ytrc_dataset_bytes: NULL");
    auto it = ctx->datasets.find(ds_id);
    if (it == ctx->datasets.end())
        return fail(YTRC_EINVAL, "ytrc_dataset_bytes: unknown dataset " + std::to_string(ds_id));
    *bytes = it->second.bytes;
    return YTRC_OK;
}

}  // extern "C"

namespace {

// Host-side preparation of one request: validation, per-brick extents,
// transfer-function tables, kernel choice and the layout of its device
// region (offsets relative to the region start).
struct PrepReq {
    ReqHost q;
    std::vector<int4> ext;
    DTF T;
    std::vector<float4> tft;
    size_t o_order = 0, o_ext = 0, o_zbuf = 0, o_init = 0, o_out = 0;
    size_t o_rv = 0, o_rh = 0, o_rs = 0, o_tf = 0, o_ppos = 0, o_pdir = 0;
    size_t bytes = 0;
    int tsh = TILE_SHIFT_DEFAULT;
};

int prep_request(ytrc_ctx *ctx, const ytrc_request &r, const std::string &tag, PrepReq &P,
                 int tsh = TILE_SHIFT_DEFAULT)
{
    P.tsh = tsh;
    ReqHost &q = P.q;
    q.r = r;
    if (r.kind != YTRC_PROJECTION && r.kind != YTRC_VOLUME_RENDER &&
        r.kind != YTRC_INTERPOLATED_PROJECTION && r.kind != YTRC_LIGHT_SOURCE)
        return fail(YTRC_ENOTIMPL, tag + "unknown sampler kind");
    if (r.lens != YTRC_LENS_PLANE_PARALLEL && r.lens != YTRC_LENS_PER_PIXEL)
        return fail(YTRC_ENOTIMPL, tag + "unknown lens");
    const bool pp = r.lens == YTRC_LENS_PER_PIXEL;
    if (pp && (r.pix_pos == nullptr || r.pix_dir == nullptr))
        return fail(YTRC_EINVAL, tag + "a per-pixel lens needs per-pixel origins and directions "
                                       "(image_samplers.pyx:89-96)");
    if (pp && (r.flags & YTRC_FLAG_ERT))
        return fail(YTRC_ENOTIMPL, tag + "early ray termination is implemented for the plane-parallel lens only");
    const bool partial = (r.flags & YTRC_FLAG_PARTIAL) != 0;
    if (partial && r.kind != YTRC_VOLUME_RENDER)
        return fail(YTRC_ENOTIMPL, tag + "partial maps are implemented for VolumeRenderSampler "
                                         "(projection partials are plain images, which add)");
    if (partial && pp)
        return fail(YTRC_ENOTIMPL, tag + "partial maps are implemented for the plane-parallel lens only");
    if (partial && r.image_in != nullptr)
        return fail(YTRC_EINVAL, tag + "a partial map starts from the identity: no image_in");
    if (r.nv[0] <= 0 || r.nv[1] <= 0) return fail(YTRC_EINVAL, tag + "image size must be > 0");
    auto dsit = ctx->datasets.find(r.ds_id);
    if (dsit == ctx->datasets.end()) return fail(YTRC_EINVAL, tag + "unknown dataset");
    const Dataset &ds = dsit->second;
    if (r.n_order < 0 || (r.n_order > 0 && r.order == nullptr))
        return fail(YTRC_EINVAL, tag + "bad brick order");
    for (int k = 0; k < r.n_order; k++) {
        if (r.order[k] < 0 || r.order[k] >= ds.n_bricks)
            return fail(YTRC_EINVAL, tag + "brick order index out of range");
        const DBrick &hb = ds.h_bricks[r.order[k]];
        if (r.kind == YTRC_PROJECTION && hb.vertex)
            return fail(YTRC_EINVAL, tag + "ProjectionSampler needs cell-centred bricks");
        if (r.kind == YTRC_VOLUME_RENDER && !hb.vertex)
            return fail(YTRC_EINVAL, tag + "VolumeRenderSampler needs vertex-centred bricks");
        if ((r.kind == YTRC_INTERPOLATED_PROJECTION || r.kind == YTRC_LIGHT_SOURCE) && !hb.vertex)
            return fail(YTRC_EINVAL, tag + "this sampler needs vertex-centred bricks");
    }
    q.npix = (int64_t)r.nv[0] * r.nv[1];
    const int th = 64 >> tsh, tw = 1 << tsh;
    const int t0 = (r.nv[0] + th - 1) / th, t1 = (r.nv[1] + tw - 1) / tw;
    // the persistent queue's items hold a tile index in 24 bits
    if ((int64_t)t0 * t1 >= (1 << 24)) return fail(YTRC_EINVAL, tag + "more than 2^24 ray tiles in one image");
    q.ntiles = t0 * t1;
    q.nfp = ds.nfp;
    const bool rec = (r.flags & YTRC_FLAG_RECORD) != 0;
    DTF &T = P.T;
    memset(&T, 0, sizeof(T));
    if (r.kind == YTRC_PROJECTION) {
        q.nch_out = std::min(4, ds.nf);
    } else if (r.kind == YTRC_INTERPOLATED_PROJECTION) {
        if (r.n_samples <= 0) return fail(YTRC_EINVAL, tag + "n_samples must be > 0");
        if (r.flags & YTRC_FLAG_ERT)
            return fail(YTRC_ENOTIMPL, tag + "early ray termination applies to volume rendering only");
        q.nch_out = std::min(3, ds.nf);  // image_samplers.pyx:399
    } else {
        if (r.tf == nullptr) return fail(YTRC_EINVAL, tag + "this sampler needs a transfer function");
        if (r.kind == YTRC_LIGHT_SOURCE && (r.flags & YTRC_FLAG_ERT))
            return fail(YTRC_ENOTIMPL, tag + "early ray termination is implemented for VolumeRenderSampler only");
        const ytrc_tf_desc &tf = *r.tf;
        if (tf.n_fits < 0 || tf.n_fits > 6)
            return fail(YTRC_EINVAL, tag + "n_fits must be <= 6 (image_samplers.pyx:430)");
        if (r.n_samples <= 0) return fail(YTRC_EINVAL, tag + "n_samples must be > 0");
        T.n_fits = tf.n_fits;
        T.grey = tf.grey_opacity == 1;
        int tot = 0;
        for (int t = 0; t < tf.n_fits; t++) {
            if (tf.nbins[t] < 2 || tf.y[t] == nullptr)
                return fail(YTRC_EINVAL, tag + "table needs nbins >= 2 and values");
            if (tf.field_ids[t] < 0 || tf.field_ids[t] >= ds.nf || tf.weight_field_ids[t] < -1 ||
                tf.weight_field_ids[t] >= ds.nf || tf.weight_table_ids[t] < -1 ||
                tf.weight_table_ids[t] >= tf.n_fits)
                return fail(YTRC_EINVAL, tag + "table field/weight ids out of range");
            T.nbins[t] = tf.nbins[t];
            T.off[t] = tot;
            T.field_id[t] = tf.field_ids[t];
            T.wfield[t] = tf.weight_field_ids[t];
            T.wtable[t] = tf.weight_table_ids[t];
            const double b0 = tf.bounds[2 * t], b1 = tf.bounds[2 * t + 1];
            T.b0[t] = (float)b0;
            T.b1[t] = (float)b1;
            // FIT_initialize_table (field_interpolation_tables.pxd:44-61) in FP64.
            const double dbin = (b1 - b0) / (double)(tf.nbins[t] - 1);
            const double idbin = 1.0 / dbin;
            T.idbin[t] = (float)idbin;
            for (int k = 0; k < tf.nbins[t]; k++) {
                float4 q4;
                q4.x = (float)tf.y[t][k];
                if (k < tf.nbins[t] - 1) {
                    q4.y = (float)((tf.y[t][k + 1] - tf.y[t][k]) * idbin);
                    q4.z = (float)(b0 + (double)k * dbin);
                } else {
                    q4.y = 0.0f;
                    q4.z = 0.0f;
                }
                q4.w = 0.0f;
                P.tft.push_back(q4);
            }
            tot += tf.nbins[t];
        }
        for (int c = 0; c < 6; c++) {
            if (tf.field_table_ids[c] < 0 || tf.field_table_ids[c] >= 6)
                return fail(YTRC_EINVAL, tag + "field_table_ids out of range");
            T.ftable[c] = tf.field_table_ids[c];
        }
        // Resolve FIT_eval_transfer's in-place weight loop
        // (field_interpolation_tables.pxd:96-99) into powers of raw values.
        int pw[6][6];
        memset(pw, 0, sizeof(pw));
        for (int t = 0; t < tf.n_fits; t++) pw[t][t] = 1;
        for (int t = 0; t < tf.n_fits; t++) {
            const int w = tf.weight_table_ids[t];
            if (w == -1) continue;
            int add[6];
            for (int u = 0; u < 6; u++) add[u] = pw[w][u];
            for (int u = 0; u < 6; u++) pw[t][u] += add[u];
        }
        for (int c = 0; c < 4; c++) {
            const int t = tf.field_table_ids[c];
            if (t >= tf.n_fits) {
                T.chexp[c] = -1;  // table never evaluated: channel value 0
                continue;
            }
            int ce = 0;
            for (int u = 0; u < 6; u++) {
                if (pw[t][u] > 3)
                    return fail(YTRC_ENOTIMPL, tag + "weight-table chain raises a table to a power > 3");
                ce |= pw[t][u] << (2 * u);
            }
            T.chexp[c] = ce;
        }
        // ColorTransferFunction pattern (transfer_functions.py:385-408):
        // four tables on one field sharing bins, rgb weighted by alpha.
        bool ctf = tf.n_fits == 4;
        for (int t = 0; ctf && t < 4; t++)
            ctf = tf.field_ids[t] == tf.field_ids[0] && tf.weight_field_ids[t] == -1 &&
                  tf.nbins[t] == tf.nbins[0] && tf.bounds[2 * t] == tf.bounds[0] &&
                  tf.bounds[2 * t + 1] == tf.bounds[1];
        ctf = ctf && T.chexp[0] == (1 | (1 << 6)) && T.chexp[1] == ((1 << 2) | (1 << 6)) &&
              T.chexp[2] == ((1 << 4) | (1 << 6)) && T.chexp[3] == (1 << 6);
        // Transmittance-form compositing is exact when every sub-sample has
        // a = dt * e <= 1: bound e by the tables' largest magnitudes and dt
        // by the longest in-cell t-interval / n_samples.
        if (!T.grey) {
            bool bounded = true;
            double emax = 0.0;
            for (int t = 0; t < tf.n_fits; t++) bounded = bounded && tf.weight_field_ids[t] == -1;
            for (int c = 0; bounded && c < 3; c++) {
                if (T.chexp[c] < 0) continue;
                double ec = 1.0;
                for (int u = 0; u < tf.n_fits; u++) {
                    const int pwr = (T.chexp[c] >> (2 * u)) & 3;
                    double m = 0.0;
                    for (int kk = 0; kk < tf.nbins[u]; kk++) m = std::max(m, std::fabs(tf.y[u][kk]));
                    for (int q2 = 0; q2 < pwr; q2++) ec *= m;
                }
                emax = std::max(emax, ec);
            }
            double tcell = 0.0;
            if (!pp) {
                for (int k2 = 0; bounded && k2 < r.n_order; k2++) {
                    const DBrick &hb = ds.h_bricks[r.order[k2]];
                    double cmin = 1.0;
                    for (int a = 0; a < 3; a++)
                        if (r.vp_dir[a] != 0.0) cmin = std::min(cmin, hb.dds[a] / std::fabs(r.vp_dir[a]));
                    tcell = std::max(tcell, cmin);
                }
            } else {
                // Any chord of a cell is at most its diagonal: the longest
                // in-cell t-interval is below diag / |v_dir| over all rays.
                double dmin = 1e300, diag = 0.0;
                for (int64_t p = 0; p < (int64_t)r.nv[0] * r.nv[1]; p++) {
                    const double *v = r.pix_dir + 3 * p;
                    dmin = std::min(dmin, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
                }
                for (int k2 = 0; k2 < r.n_order; k2++) {
                    const DBrick &hb = ds.h_bricks[r.order[k2]];
                    diag = std::max(diag, std::sqrt(hb.dds[0] * hb.dds[0] + hb.dds[1] * hb.dds[1] +
                                                    hb.dds[2] * hb.dds[2]));
                }
                tcell = dmin > 0.0 ? diag / dmin : 1e300;
                bounded = bounded && std::isfinite(tcell);
            }
            q.tform = bounded && (emax * (tcell / (double)r.n_samples) * 1.001 <= 1.0);
        }
        // ColorTransferFunction pattern (transfer_functions.py:385-408) in
        // transmittance form gets its own kernel path (MODE bit 0); grey or
        // unbounded CTFs take the general per-table path.
        ctf = ctf && q.tform && r.kind == YTRC_VOLUME_RENDER;
        if (ctf) {
            // Transmittance form: per bin, the two lerps whose product is
            // the emission, e_c = (y_c + f dy_c)(y_a + f dy_a), c = r, g, b
            // (FIT_get_value's lerp, field_interpolation_tables.pxd:63-82,
            // times the alpha weight table, transfer_functions.py:385-408),
            // dy = y[b+1] - y[b], f = u - k the in-bin coordinate of entry
            // k.  Stored as {y - k dy} and {dy} for the four tables, so that
            // a lerp is one FMA in the bin coordinate u: y + f dy =
            // (y - k dy) + u dy; dy is rounded to FP32 first and the offset
            // formed from it in FP64.  Entry k = 1..nb-1 holds bin k-1;
            // entries 0, nb and nb+1 are zero (values outside (b0, b1)).
            // Entries are 48 B apart (a zero float4 of padding): with an odd
            // number of 16-B units per entry, the 16-B reads of different
            // entries start on all 16 bank groups of the LDS rather than
            // every other one, which halves the bank conflicts of the
            // scattered per-lane reads.
            const int nb = tf.nbins[0];
            const double b0 = tf.bounds[0], b1 = tf.bounds[1];
            const double idbin = 1.0 / ((b1 - b0) / (double)(nb - 1));
            std::vector<float4> lay((size_t)3 * (nb + 2), make_float4(0.f, 0.f, 0.f, 0.f));
            for (int k = 1; k < nb; k++) {
                const int bb = k - 1;
                float y[4], d[4];
                for (int t = 0; t < 4; t++) {
                    d[t] = (float)(tf.y[t][bb + 1] - tf.y[t][bb]);
                    y[t] = (float)(tf.y[t][bb] - (double)k * (double)d[t]);
                }
                lay[3 * k] = make_float4(y[0], y[1], y[2], y[3]);
                lay[3 * k + 1] = make_float4(d[0], d[1], d[2], d[3]);
            }
            P.tft = std::move(lay);
            T.idbin[0] = (float)idbin;
            T.qc1 = (float)(1.0 - b0 * idbin);
            T.qzb = (float)(nb + 1);
        }
        if (ctf) {
            tot = (int)P.tft.size();
            T.mode = 1;
        }
        q.mode = T.mode | (r.n_samples == 10 ? 2 : 0) |
                 ((r.flags & (YTRC_FLAG_ERT | YTRC_FLAG_PARTIAL)) ? 4 : 0) | (pp ? 8 : 0);
        T.total_bins = tot;
        q.lds = (size_t)tot * sizeof(float4);
        if (q.lds > ctx->max_lds)
            return fail(YTRC_EINVAL, tag + "transfer-function tables need " + std::to_string(q.lds) +
                                         " B of LDS, more than the " + std::to_string(ctx->max_lds) +
                                         " B a workgroup may allocate on this device");
        q.nch_out = (T.grey && r.kind == YTRC_VOLUME_RENDER) ? 4 : 3;
        if (partial) q.nch_out = 8;  // C[0..3], T[0..3]
    }
    if (r.kind != YTRC_VOLUME_RENDER) q.mode = pp ? 8 : 0;
    if (r.kind == YTRC_LIGHT_SOURCE) q.tform = 0;
    q.kfn = pick_kernel(r.kind, q.nfp, rec, q.mode);
    if (q.kfn == nullptr) return fail(YTRC_ENOTIMPL, tag + "no kernel for this field count");
    if (q.lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)q.kfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)q.lds);
        if (e != hipSuccess)
            return fail(YTRC_EDEVICE, tag + "hipFuncSetAttribute(LDS " + std::to_string(q.lds) + " B): " +
                                          hipGetErrorString(e));
    }
    if (r.image_in != nullptr) q.image_in.assign(r.image_in, r.image_in + q.npix * 4);
    P.ext.resize(std::max(1, r.n_order));
    for (int k = 0; k < r.n_order; k++) {
        int64_t rv[4];
        if (!pp) {
            extent_plane_parallel(r, ds.h_bricks[r.order[k]], rv);
        } else if (r.camera_data != nullptr) {
            extent_perspective(r, ds.h_bricks[r.order[k]], rv);
        } else {
            rv[0] = 0;  // calculate_extent_null (lenses.pyx:169-175); the
            rv[1] = r.nv[0];  // +-1 clip leaves the whole image
            rv[2] = 0;
            rv[3] = r.nv[1];
        }
        P.ext[k] = make_int4((int)rv[0], (int)rv[1], (int)rv[2], (int)rv[3]);
    }
    // Device region layout.
    size_t off = 0;
    P.o_order = off;
    off = align_up(off + sizeof(int32_t) * std::max(1, r.n_order), 256);
    P.o_ext = off;
    off = align_up(off + sizeof(int4) * P.ext.size(), 256);
    P.o_zbuf = off;
    if (r.zbuffer) off = align_up(off + sizeof(double) * q.npix, 256);
    P.o_init = off;
    if (r.image_in && !additive_kind(r.kind)) off = align_up(off + sizeof(float) * 4 * q.npix, 256);
    P.o_out = off;
    off = align_up(off + sizeof(float) * std::max(4, q.nch_out) * q.npix, 256);
    P.o_rv = P.o_rh = P.o_rs = off;
    if (rec) {
        P.o_rv = off;
        off = align_up(off + 8 * q.npix, 256);
        P.o_rh = off;
        off = align_up(off + 8 * q.npix, 256);
        P.o_rs = off;
        off = align_up(off + 8 * q.npix, 256);
    }
    P.o_tf = off;
    if (has_tf(r.kind)) off = align_up(off + sizeof(DTF), 256);
    P.o_ppos = P.o_pdir = off;
    if (pp) {
        P.o_ppos = off;
        off = align_up(off + sizeof(double) * 3 * q.npix, 256);
        P.o_pdir = off;
        off = align_up(off + sizeof(double) * 3 * q.npix, 256);
    }
    P.bytes = off;
    return YTRC_OK;
}

// Copies a prepared request into its device region `A` and fills its
// device descriptor.  `tf_table` is the device address of its (deduplicated)
// transfer-function table, at float4 offset `tf_slot` of the launch's table.
int stage_request(ytrc_ctx *ctx, PrepReq &P, char *A, const float4 *tf_table, int tf_slot, DImg &d,
                  std::vector<float> &stage)
{
    const ytrc_request &r = P.q.r;
    const ReqHost &q = P.q;
    const Dataset &ds = ctx->datasets[r.ds_id];
    const bool rec = (r.flags & YTRC_FLAG_RECORD) != 0;
    memset(&d, 0, sizeof(d));
    memcpy(d.vp_pos, r.vp_pos, sizeof(d.vp_pos));
    memcpy(d.vp_dir, r.vp_dir, sizeof(d.vp_dir));
    memcpy(d.width, r.width, sizeof(d.width));
    d.bricks = ds.d_bricks;
    d.pool = ds.d_pool;
    d.masks = ds.d_masks;
    d.order = (const int32_t *)(A + P.o_order);
    d.ext = (const int4 *)(A + P.o_ext);
    d.zbuf = r.zbuffer ? (const double *)(A + P.o_zbuf) : nullptr;
    const bool init_dev = r.image_in != nullptr && !additive_kind(r.kind);
    d.init = init_dev ? (const float *)(A + P.o_init) : nullptr;
    d.out = (float *)(A + P.o_out);
    d.rec_visits = rec ? (int64_t *)(A + P.o_rv) : nullptr;
    d.rec_hash = rec ? (uint64_t *)(A + P.o_rh) : nullptr;
    d.rec_samples = rec ? (int64_t *)(A + P.o_rs) : nullptr;
    d.tf = has_tf(r.kind) ? (const DTF *)(A + P.o_tf) : nullptr;
    for (int i = 0; i < 3; i++) d.ldir[i] = r.light_dir[i];
    for (int i = 0; i < 4; i++) d.lrgba[i] = r.light_rgba[i];
    d.nv0 = r.nv[0];
    d.nv1 = r.nv[1];
    d.kind = r.kind;
    d.n_samples = r.n_samples;
    d.n_order = r.n_order;
    d.nf = ds.nf;
    d.nfp = ds.nfp;
    d.tsh = P.tsh;
    d.tiles0 = (r.nv[0] + (64 >> P.tsh) - 1) / (64 >> P.tsh);
    d.tiles1 = (r.nv[1] + (1 << P.tsh) - 1) / (1 << P.tsh);
    d.flags = (int32_t)r.flags;
    d.nch_out = q.nch_out;
    d.inv_ns = r.n_samples > 0 ? 1.0 / (double)r.n_samples : 0.0;
    d.tform = q.tform;
    if (r.n_order > 0) {
        HIPCHK(hipMemcpy(A + P.o_order, r.order, sizeof(int32_t) * r.n_order, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(A + P.o_ext, P.ext.data(), sizeof(int4) * r.n_order, hipMemcpyHostToDevice));
    }
    if (r.zbuffer)
        HIPCHK(hipMemcpy(A + P.o_zbuf, r.zbuffer, sizeof(double) * q.npix, hipMemcpyHostToDevice));
    if (r.lens == YTRC_LENS_PER_PIXEL) {
        d.pix_pos = (const double *)(A + P.o_ppos);
        d.pix_dir = (const double *)(A + P.o_pdir);
        HIPCHK(hipMemcpy(A + P.o_ppos, r.pix_pos, sizeof(double) * 3 * q.npix, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(A + P.o_pdir, r.pix_dir, sizeof(double) * 3 * q.npix, hipMemcpyHostToDevice));
    }
    if (init_dev) {
        stage.resize((size_t)q.npix * 4);
        for (int64_t p = 0; p < q.npix; p++)
            for (int c = 0; c < 4; c++) stage[(size_t)c * q.npix + p] = (float)r.image_in[p * 4 + c];
        HIPCHK(hipMemcpy(A + P.o_init, stage.data(), sizeof(float) * 4 * q.npix, hipMemcpyHostToDevice));
    }
    if (has_tf(r.kind)) {
        DTF T = P.T;
        T.table = tf_table + tf_slot;
        d.tf_slot = tf_slot;
        HIPCHK(hipMemcpy(A + P.o_tf, &T, sizeof(DTF), hipMemcpyHostToDevice));
    }
    return YTRC_OK;
}

// The float64 value of channel value `o` (float32, from the device) of a
// pixel whose input was `in` (image_samplers.pyx:177-196 keeps the image in
// float64 across bricks):
//   projection samplers: the device's sum starts from 0, so in + o, and a
//   pixel no visit touched (o == 0) keeps `in` exactly;
//   compositing samplers: the device started from float(in); where the
//   result still equals float(in), nothing changed the pixel by a float32
//   step, and `in` is returned exactly (for an untouched pixel, or one whose
//   samples emit nothing, yt leaves the float64 value as it was).
static inline double merge_input(int32_t kind, float o, double in)
{
    if (additive_kind(kind)) return o == 0.0f ? in : in + (double)o;
    return o == (float)in ? in : (double)o;
}

// Copies a finished request's output (and image_in for channels it does
// not write) into the caller's float64 (nv0, nv1, 4) image.
int read_request_image(const ReqHost &q, const DImg &d, double *image_out)
{
    if (q.r.flags & YTRC_FLAG_PARTIAL)
        return fail(YTRC_EINVAL, "a YTRC_FLAG_PARTIAL request has a map, not an image: "
                                 "read it with ytrc_batch_read_partial");
    std::vector<float> h((size_t)q.npix * q.nch_out);
    HIPCHK(hipMemcpy(h.data(), d.out, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
    const bool has_in = !q.image_in.empty();
    for (int64_t p = 0; p < q.npix; p++) {
        for (int c = 0; c < 4; c++) {
            const double in = has_in ? q.image_in[(size_t)p * 4 + c] : 0.0;
            double v;
            if (c < q.nch_out)
                v = has_in ? merge_input(q.r.kind, h[(size_t)c * q.npix + p], in)
                           : (double)h[(size_t)c * q.npix + p];
            else
                v = in;
            image_out[(size_t)p * 4 + c] = v;
        }
    }
    return YTRC_OK;
}

}  // namespace

extern "C" {

int ytrc_batch_create(ytrc_ctx *ctx, int32_t n_req, const ytrc_request *reqs,
                      uint32_t batch_flags, ytrc_batch **out)
{
    if (ctx == nullptr || reqs == nullptr || out == nullptr)
        return fail(YTRC_EINVAL, "ytrc_batch_create: NULL argument");
    if (n_req <= 0) return fail(YTRC_EINVAL, "ytrc_batch_create: n_req must be > 0");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<PrepReq> P(n_req);
    size_t off = 0;
    std::vector<size_t> base(n_req);
    const int tcode = (int)((batch_flags >> 4) & 3);
    if (tcode == 3) return fail(YTRC_EINVAL, "ytrc_batch_create: unknown tile shape");
    const int tsh = tcode == 1 ? 4 : (tcode == 2 ? 2 : TILE_SHIFT_DEFAULT);
    const int per_item = 1 << (int)((batch_flags >> 6) & 3);  // tiles per work item
    for (int i = 0; i < n_req; i++) {
        const int rc = prep_request(ctx, reqs[i], "request " + std::to_string(i) + ": ", P[i], tsh);
        if (rc != YTRC_OK) return rc;
        base[i] = off;
        off = align_up(off + P[i].bytes, 256);
    }
    ytrc_batch *B = new ytrc_batch();
    B->ctx = ctx;
    B->flags = batch_flags;
    B->reqs.resize(n_req);
    B->h_imgs.resize(n_req);
    for (int i = 0; i < n_req; i++) B->reqs[i] = P[i].q;
    auto bad = [&](int code, const std::string &m) {
        if (B->arena) {
            hipFree(B->arena);
            ctx->used -= B->arena_bytes;
        }
        if (B->h_stop) hipHostFree(B->h_stop);
        if (B->ev0) hipEventDestroy(B->ev0);
        if (B->ev1) hipEventDestroy(B->ev1);
        if (B->ev_fork) hipEventDestroy(B->ev_fork);
        if (B->ev_join) hipEventDestroy(B->ev_join);
        for (auto &ev : B->ev_req)
            if (ev) hipEventDestroy(ev);
        delete B;
        return fail(code, m);
    };
    // Persistent launches: one per kernel variant, items = (request, tile).
    // A launch's workgroups stage its requests' distinct transfer-function
    // tables in LDS, so with several tables per variant the batch flag
    // YTRC_BATCH_TF_LDS_* caps the table bytes of one launch: the variant's
    // requests are then split over launches by table (a table is never
    // split; one larger than the cap gets a launch of its own).  Fewer LDS
    // bytes per workgroup let more workgroups share a CU.
    const int lcode = (int)((batch_flags >> 8) & 3);
    const size_t tf_cap = lcode == 1 ? (32u << 10) : (lcode == 2 ? (16u << 10) : (lcode == 3 ? 1u : 0u));
    // YTRC_BATCH_VERTEX_LDS: the volume renders that take the ColorTransferFunction
    // fast walk with 10 sub-samples (mode 3, one field, no record) run the
    // variant that stages vertex blocks in LDS (mode bit 4).
    const bool vlds = (batch_flags & YTRC_BATCH_VERTEX_LDS) != 0;
    std::map<std::tuple<int, int, int, int>, std::vector<int>> gmap;
    for (int i = 0; i < n_req; i++) {
        const bool rec = (reqs[i].flags & YTRC_FLAG_RECORD) != 0;
        int mode = P[i].q.mode;
        if (vlds && reqs[i].kind == YTRC_VOLUME_RENDER && P[i].q.nfp == 1 && !rec && mode == 3) mode |= 16;
        gmap[std::make_tuple(reqs[i].kind, P[i].q.nfp, rec ? 1 : 0, mode)].push_back(i);
    }
    struct LaunchPlan {
        std::tuple<int, int, int, int> variant;
        std::vector<int> ids;
        std::vector<std::string> keys;  // its distinct tables, in first-use order
        size_t table_bytes = 0;
    };
    std::vector<LaunchPlan> plans;
    for (auto &kv : gmap) {
        const bool tfk = has_tf(std::get<0>(kv.first));
        std::vector<LaunchPlan> mine;
        std::map<std::string, int> plan_of;  // table -> index in mine
        for (int i : kv.second) {
            std::string key;
            size_t tb = 0;
            if (tfk) {
                key.assign((const char *)P[i].tft.data(), P[i].tft.size() * sizeof(float4));
                tb = key.size();
            }
            auto it = plan_of.find(key);
            if (it == plan_of.end()) {
                int pi = -1;
                if (!mine.empty() && (tf_cap == 0 || mine.back().keys.empty() ||
                                      mine.back().table_bytes + tb <= tf_cap))
                    pi = (int)mine.size() - 1;
                if (pi < 0) {
                    mine.emplace_back();
                    mine.back().variant = kv.first;
                    pi = (int)mine.size() - 1;
                }
                if (tfk) {
                    mine[pi].keys.push_back(key);
                    mine[pi].table_bytes += tb;
                }
                it = plan_of.emplace(key, pi).first;
            }
            mine[it->second].ids.push_back(i);
        }
        for (auto &m : mine) plans.push_back(std::move(m));
    }
    // The batch table: each launch's distinct tables, launch after launch.
    std::vector<float4> tf_all;
    std::vector<int> slot(n_req, 0), rel_slot(n_req, 0);
    std::vector<int> plan_base(plans.size(), 0), plan_bins(plans.size(), 0);
    for (size_t L = 0; L < plans.size(); L++) {
        const int b0 = (int)tf_all.size();
        plan_base[L] = b0;
        std::map<std::string, int> at;
        for (const std::string &k : plans[L].keys) {
            at[k] = (int)tf_all.size();
            const float4 *q = (const float4 *)k.data();
            tf_all.insert(tf_all.end(), q, q + k.size() / sizeof(float4));
        }
        plan_bins[L] = (int)tf_all.size() - b0;
        if (!plans[L].keys.empty())
            for (int i : plans[L].ids) {
                const std::string key((const char *)P[i].tft.data(), P[i].tft.size() * sizeof(float4));
                slot[i] = at[key];
                rel_slot[i] = slot[i] - b0;
            }
    }
    const size_t tfall_off = off;
    off = align_up(off + sizeof(float4) * std::max<size_t>(1, tf_all.size()), 256);
    // Each image's tiles are cut into QSPLIT contiguous chunks and chunk r
    // goes to queue range r (the range one XCD drains first), so every XCD
    // renders a compact part of every image; within a range the images
    // advance together (items ordered by the fraction of their chunk done,
    // larger images first on ties).  Waves of small images, whose rays are
    // sparser than the cells and wait on memory, then run beside waves of
    // large images, which keep the VALUs busy: 5-6 % faster than running
    // the images one after another, larger first.
    std::vector<std::vector<int2>> gitems;
    std::vector<size_t> gitems_off;
    for (size_t L = 0; L < plans.size(); L++) {
        std::vector<int> ids = plans[L].ids;
        std::stable_sort(ids.begin(), ids.end(),
                         [&](int a, int b) { return P[a].q.ntiles > P[b].q.ntiles; });
        std::vector<int2> items;
        struct Key {
            double frac;
            int ord;
            int2 item;
        };
        // Single-tile items take each image's tiles in Z order, so that the
        // tiles in flight in a range form compact patches, which share more
        // of their bricks' lines in L2: on the 512^3 batch the projection
        // kernel 1.2 % faster, the volume-render kernel alone 0.4 % slower,
        // the two launches together 0.7 % faster than in row order.
        const bool z = per_item == 1;
        std::vector<std::vector<int>> zord(ids.size());
        if (z)
            for (size_t o = 0; o < ids.size(); o++) {
                const PrepReq &Q = P[ids[o]];
                const int th = 64 >> Q.tsh, tw = 1 << Q.tsh;
                zord[o] = tiles_z_order((Q.q.r.nv[0] + th - 1) / th, (Q.q.r.nv[1] + tw - 1) / tw);
            }
        for (int r = 0; r < QSPLIT; r++) {
            std::vector<Key> ks;
            for (size_t o = 0; o < ids.size(); o++) {
                const int i = ids[o];
                const int T = P[i].q.ntiles;
                const int t0 = (int)((int64_t)r * T / QSPLIT), t1 = (int)((int64_t)(r + 1) * T / QSPLIT);
                // an item: up to per_item consecutive tiles, the count in
                // the top 8 bits of its tile field
                for (int t = t0; t < t1; t += per_item) {
                    const int cnt = std::min(per_item, t1 - t);
                    const int tile = z ? zord[o][(size_t)t] : t;
                    ks.push_back({(double)(t - t0) / (double)(t1 - t0), (int)o, make_int2(i, tile | (cnt << 24))});
                }
            }
            std::stable_sort(ks.begin(), ks.end(), [](const Key &a, const Key &b) {
                return a.frac < b.frac || (a.frac == b.frac && a.ord < b.ord);
            });
            for (const Key &k : ks) items.push_back(k.item);
        }
        PGroup g;
        g.kind = std::get<0>(plans[L].variant);
        g.nfp = std::get<1>(plans[L].variant);
        g.rec = std::get<2>(plans[L].variant) != 0;
        g.mode = std::get<3>(plans[L].variant);
        g.fn = pick_persist(g.kind, g.nfp, g.rec, g.mode);
        g.n_items = (int)items.size();
        g.counter = (int)B->groups.size();
        g.tf_base = plan_base[L];
        g.tf_bins = plan_bins[L];
        g.lds = has_tf(g.kind) ? sizeof(float4) * (size_t)g.tf_bins : 0;
        // up to 4 waves' buffers follow the tables
        const size_t lds_most = g.lds + 4 * sizeof(unsigned) * required_wave_words(g.kind, g.nfp, g.rec, g.mode);
        if (lds_most > ctx->max_lds)
            return bad(YTRC_EINVAL, "ytrc_batch_create: the distinct transfer functions of one launch need " +
                                        std::to_string(lds_most) + " B of LDS, more than the " +
                                        std::to_string(ctx->max_lds) + " B a workgroup may allocate on this device");
        if (lds_most > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute((const void *)g.fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)lds_most);
            if (e != hipSuccess)
                return bad(YTRC_EDEVICE, std::string("ytrc_batch_create: hipFuncSetAttribute: ") + hipGetErrorString(e));
        }
        B->groups.push_back(g);
        gitems_off.push_back(off);
        off = align_up(off + sizeof(int2) * std::max<size_t>(1, items.size()), 256);
        gitems.push_back(std::move(items));
    }
    const size_t counters_off = off;
    off = align_up(off + sizeof(unsigned int) * QCOUNTERS * std::max<size_t>(1, B->groups.size()), 256);
    const size_t imgs_off = off;
    off = align_up(off + sizeof(DImg) * n_req, 256);
    if (ctx->budget != 0 && ctx->used + off > ctx->budget)
        return bad(YTRC_EBUDGET, "ytrc_batch_create: " + std::to_string(off) + " bytes exceed the HBM budget");
    hipError_t e = hipMalloc(&B->arena, off);
    if (e != hipSuccess) return bad(YTRC_EDEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    B->arena_bytes = off;
    ctx->used += off;
    char *A = B->arena;
    std::vector<float> stage;
    for (int i = 0; i < n_req; i++) {
        const int rc = stage_request(ctx, P[i], A + base[i], (const float4 *)(A + tfall_off), slot[i],
                                     B->h_imgs[i], stage);
        if (rc != YTRC_OK) return bad(rc, g_err);
        // the persistent kernels read the table from their launch's block
        B->h_imgs[i].tf_slot = rel_slot[i];
    }
    B->tf_all = (const float4 *)(A + tfall_off);
    B->tf_bins = (int)tf_all.size();
    if (!tf_all.empty()) {
        e = hipMemcpy(A + tfall_off, tf_all.data(), sizeof(float4) * tf_all.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return bad(YTRC_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    for (size_t g = 0; g < B->groups.size(); g++) {
        B->groups[g].items = (const int2 *)(A + gitems_off[g]);
        if (!gitems[g].empty()) {
            e = hipMemcpy(A + gitems_off[g], gitems[g].data(), sizeof(int2) * gitems[g].size(),
                          hipMemcpyHostToDevice);
            if (e != hipSuccess) return bad(YTRC_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
        }
    }
    B->counters = (unsigned int *)(A + counters_off);
    B->d_imgs = (DImg *)(A + imgs_off);
    // The stop flag: fine-grained pinned host memory that the kernels read
    // at system scope once per tile.
    e = hipHostMalloc((void **)&B->h_stop, sizeof(int), hipHostMallocCoherent | hipHostMallocMapped);
    if (e != hipSuccess) return bad(YTRC_EDEVICE, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    __atomic_store_n(B->h_stop, 0, __ATOMIC_SEQ_CST);
    int *d_stop = nullptr;
    e = hipHostGetDevicePointer((void **)&d_stop, B->h_stop, 0);
    if (e != hipSuccess) return bad(YTRC_EDEVICE, std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
    for (int i = 0; i < n_req; i++) B->h_imgs[i].stop = d_stop;
    e = hipMemcpy(B->d_imgs, B->h_imgs.data(), sizeof(DImg) * n_req, hipMemcpyHostToDevice);
    if (e != hipSuccess) return bad(YTRC_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
    if ((e = hipEventCreate(&B->ev0)) != hipSuccess || (e = hipEventCreate(&B->ev1)) != hipSuccess)
        return bad(YTRC_EDEVICE, std::string("hipEventCreate: ") + hipGetErrorString(e));
    if (batch_flags & YTRC_BATCH_PERSISTENT) {
        if ((e = hipEventCreateWithFlags(&B->ev_fork, hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&B->ev_join, hipEventDisableTiming)) != hipSuccess)
            return bad(YTRC_EDEVICE, std::string("hipEventCreate: ") + hipGetErrorString(e));
    } else {
        B->ev_req.assign(2 * n_req, nullptr);
        for (auto &ev : B->ev_req)
            if ((e = hipEventCreate(&ev)) != hipSuccess)
                return bad(YTRC_EDEVICE, std::string("hipEventCreate: ") + hipGetErrorString(e));
    }
    // Hold the datasets until the batch is destroyed (ytrc_dataset_free
    // refuses a dataset in use).
    for (int i = 0; i < n_req; i++)
        if (std::find(B->ds_ids.begin(), B->ds_ids.end(), reqs[i].ds_id) == B->ds_ids.end())
            B->ds_ids.push_back(reqs[i].ds_id);
    for (int id : B->ds_ids) ctx->datasets[id].refs++;
    *out = B;
    return YTRC_OK;
}

int ytrc_batch_launch_plan(const ytrc_batch *b, int32_t *n_launches, uint64_t *max_tf_bytes)
{
    if (b == nullptr || n_launches == nullptr || max_tf_bytes == nullptr)
        return fail(YTRC_EINVAL, "ytrc_batch_launch_plan: NULL argument");
    int n = 0;
    uint64_t mx = 0;
    if (b->flags & YTRC_BATCH_PERSISTENT)
        for (const PGroup &g : b->groups)
            if (g.n_items > 0) {
                n++;
                mx = std::max<uint64_t>(mx, g.lds);
            }
    *n_launches = n;
    *max_tf_bytes = mx;
    return YTRC_OK;
}

int ytrc_batch_set_launch(ytrc_batch *b, int32_t waves_per_group, int32_t groups_per_cu)
{
    if (b == nullptr) return fail(YTRC_EINVAL, "ytrc_batch_set_launch: NULL batch");
    if (waves_per_group != 1 && waves_per_group != 2 && waves_per_group != 4)
        return fail(YTRC_EINVAL, "ytrc_batch_set_launch: waves_per_group must be 1, 2 or 4");
    if (groups_per_cu < 0 || groups_per_cu > 64)
        return fail(YTRC_EINVAL, "ytrc_batch_set_launch: groups_per_cu must be in 0..64");
    b->waves_per_group = waves_per_group;
    b->groups_per_cu = groups_per_cu;
    if (b->gexec) {
        hipGraphExecDestroy(b->gexec);
        hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
    }
    return YTRC_OK;
}

namespace {

int enqueue_persistent(ytrc_batch *b, hipStream_t s)
{
    const int ng = (int)b->groups.size();
    hipLaunchKernelGGL(k_zero_counters, dim3((ng * QCOUNTERS + 255) / 256), dim3(256), 0, s, b->counters,
                       ng * QCOUNTERS);
    HIPCHK(hipGetLastError());
    const int threads = 64 * b->waves_per_group;
    // The launches run concurrently: the one with the most estimated work
    // (tiles x a per-tile cost by kind) first on `s`, the others after it
    // on the second stream, joined back to `s`.  The first launch's
    // workgroups take the GPU but one per CU; the others' start in that room
    // and fill the GPU as the first one's waves finish, so its tail (the
    // longest) is not left idle.
    std::vector<int> order;
    for (int i = 0; i < (int)b->groups.size(); i++)
        if (b->groups[i].n_items > 0) order.push_back(i);
    auto work = [&](int i) {
        const PGroup &g = b->groups[i];
        const double per_tile = g.kind == YTRC_PROJECTION ? 1.0 : (g.kind == YTRC_VOLUME_RENDER ? 5.0 : 4.0);
        return per_tile * (double)g.n_items;
    };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) > work(y); });
    const bool fork = order.size() >= 2;
    if (fork) {
        HIPCHK(hipEventRecord(b->ev_fork, s));
        HIPCHK(hipStreamWaitEvent(b->ctx->stream2, b->ev_fork, 0));
    }
    for (size_t k = 0; k < order.size(); k++) {
        PGroup &g = b->groups[order[k]];
        hipStream_t gs = (k == 0) ? s : b->ctx->stream2;
        // each wave's buffer after the tables
        const int ww = required_wave_words(g.kind, g.nfp, g.rec, g.mode);
        const size_t lds = g.lds + (size_t)b->waves_per_group * ww * sizeof(unsigned);
        int per_cu = b->groups_per_cu;
        if (per_cu == 0) {
            int nb = 0;
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, g.fn, threads, lds));
            per_cu = std::max(1, (fork && k == 0) ? nb - 1 : nb);
        }
        g.grid = std::max(1, std::min(per_cu * b->ctx->n_cu,
                                      (g.n_items + b->waves_per_group - 1) / b->waves_per_group));
        hipLaunchKernelGGL(g.fn, dim3(g.grid), dim3(threads), lds, gs, (const DImg *)b->d_imgs,
                           g.items, g.n_items, b->counters + (size_t)g.counter * QCOUNTERS, b->tf_all + g.tf_base,
                           g.tf_bins, ww);
        HIPCHK(hipGetLastError());
    }
    if (fork) {
        HIPCHK(hipEventRecord(b->ev_join, b->ctx->stream2));
        HIPCHK(hipStreamWaitEvent(s, b->ev_join, 0));
    }
    return YTRC_OK;
}

int enqueue_plain(ytrc_batch *b, hipStream_t s, bool events)
{
    const int wpg = b->waves_per_group;
    for (size_t i = 0; i < b->reqs.size(); i++) {
        const ReqHost &q = b->reqs[i];
        const int groups = (q.ntiles + wpg - 1) / wpg;
        if (events) HIPCHK(hipEventRecord(b->ev_req[2 * i], s));
        if (groups > 0) {
            hipLaunchKernelGGL(q.kfn, dim3(groups), dim3(64 * wpg), q.lds, s,
                               (const DImg *)b->d_imgs, (int)i);
            HIPCHK(hipGetLastError());
        }
        if (events) HIPCHK(hipEventRecord(b->ev_req[2 * i + 1], s));
    }
    return YTRC_OK;
}

}  // namespace

int ytrc_batch_launch(ytrc_batch *b)
{
    if (b == nullptr) return fail(YTRC_EINVAL, "ytrc_batch_launch: NULL batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    // a new launch clears an earlier stop request
    __atomic_store_n(b->h_stop, 0, __ATOMIC_SEQ_CST);
    b->cancelled = false;
    hipStream_t s = b->ctx->stream;
    const bool persistent = (b->flags & YTRC_BATCH_PERSISTENT) != 0;
    if ((b->flags & YTRC_BATCH_GRAPH) && b->gexec == nullptr) {
        // Capture the launch sequence once; replays reuse it.
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        int rc = persistent ? enqueue_persistent(b, s) : enqueue_plain(b, s, false);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc != YTRC_OK) return rc;
        if (e != hipSuccess) return fail(YTRC_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    }
    HIPCHK(hipEventRecord(b->ev0, s));
    int rc;
    if (b->gexec) {
        HIPCHK(hipGraphLaunch(b->gexec, s));
        rc = YTRC_OK;
    } else {
        rc = persistent ? enqueue_persistent(b, s) : enqueue_plain(b, s, true);
    }
    if (rc != YTRC_OK) return rc;
    HIPCHK(hipEventRecord(b->ev1, s));
    b->launched += 1;
    return YTRC_OK;
}

int ytrc_render_noalloc(ytrc_ctx *ctx, const ytrc_request *req, double *image_out)
{
    if (ctx == nullptr || req == nullptr || image_out == nullptr)
        return fail(YTRC_EINVAL, "ytrc_render_noalloc: NULL argument");
    if (req->flags & (YTRC_FLAG_RECORD | YTRC_FLAG_PARTIAL))
        return fail(YTRC_EINVAL, "ytrc_render_noalloc: record and partial-map requests need the batch "
                                 "API (their extra outputs are not staged in the fixed array)");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->scratch == nullptr || ctx->scratch_mu == nullptr)
        return fail(YTRC_EDEVICE, "ytrc_render_noalloc: no scratch address for this device");
    std::lock_guard<std::mutex> sl(*ctx->scratch_mu);
    HIPCHK(hipSetDevice(ctx->device));
    PrepReq P;
    int rc = prep_request(ctx, *req, "ytrc_render_noalloc: ", P, TILE_SHIFT_DEFAULT);
    if (rc != YTRC_OK) return rc;
    // Layout in the array: the request's region (as in a batch), its
    // transfer-function table, its tile list, the queue counters and its
    // descriptor.
    size_t off = align_up(P.bytes, 256);
    const size_t tf_off = off;
    off = align_up(off + sizeof(float4) * std::max<size_t>(1, P.tft.size()), 256);
    const size_t items_off = off;
    off = align_up(off + sizeof(int2) * (size_t)std::max(1, P.q.ntiles), 256);
    const size_t cnt_off = off;
    off = align_up(off + sizeof(unsigned int) * QCOUNTERS, 256);
    const size_t img_off = off;
    off = align_up(off + sizeof(DImg), 256);
    if (off > (size_t)YTRC_NOALLOC_SCRATCH_BYTES)
        return fail(YTRC_EINVAL, "ytrc_render_noalloc: the request needs " + std::to_string(off) +
                                     " B of device staging, more than the " +
                                     std::to_string((size_t)YTRC_NOALLOC_SCRATCH_BYTES) +
                                     " B fixed array; use ytrc_batch_create");
    char *S = ctx->scratch;
    DImg d;
    std::vector<float> stage;
    rc = stage_request(ctx, P, S, (const float4 *)(S + tf_off), 0, d, stage);
    if (rc != YTRC_OK) return rc;
    if (!P.tft.empty())
        HIPCHK(hipMemcpy(S + tf_off, P.tft.data(), sizeof(float4) * P.tft.size(), hipMemcpyHostToDevice));
    // One request: the tiles in order (the QSPLIT queue ranges are its
    // QSPLIT contiguous chunks, as in a persistent batch of this request).
    std::vector<int2> items((size_t)P.q.ntiles);
    for (int t = 0; t < P.q.ntiles; t++) items[(size_t)t] = make_int2(0, t);
    if (!items.empty())
        HIPCHK(hipMemcpy(S + items_off, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(S + img_off, &d, sizeof(DImg), hipMemcpyHostToDevice));
    const pfn_t fn = pick_persist(req->kind, P.q.nfp, false, P.q.mode);
    if (fn == nullptr) return fail(YTRC_ENOTIMPL, "ytrc_render_noalloc: no kernel for this request");
    const int wpg = 4, threads = 64 * wpg;
    const int ww = required_wave_words(req->kind, P.q.nfp, false, P.q.mode);
    const size_t lds = (has_tf(req->kind) ? sizeof(float4) * P.tft.size() : 0) + (size_t)wpg * ww * sizeof(unsigned);
    if (lds > 64 * 1024)
        HIPCHK(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipStream_t s = ctx->stream;
    unsigned int *counters = (unsigned int *)(S + cnt_off);
    // the array keeps the previous call's counters: clear them
    hipLaunchKernelGGL(k_zero_counters, dim3((QCOUNTERS + 255) / 256), dim3(256), 0, s, counters, QCOUNTERS);
    HIPCHK(hipGetLastError());
    if (P.q.ntiles > 0) {
        int nb = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, threads, lds));
        const int grid = std::max(1, std::min(std::max(1, nb) * ctx->n_cu, (P.q.ntiles + wpg - 1) / wpg));
        hipLaunchKernelGGL(fn, dim3(grid), dim3(threads), lds, s, (const DImg *)(S + img_off),
                           (const int2 *)(S + items_off), P.q.ntiles, counters,
                           (const float4 *)(S + tf_off), (int)P.tft.size(), ww);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s));
    return read_request_image(P.q, d, image_out);
}

int ytrc_batch_request_ms(ytrc_batch *b, int32_t i, float *ms)
{
    if (b == nullptr || ms == nullptr) return fail(YTRC_EINVAL, "ytrc_batch_request_ms: NULL");
    if (i < 0 || i >= (int)b->reqs.size()) return fail(YTRC_EINVAL, "ytrc_batch_request_ms: bad index");
    if (b->ev_req.empty() || b->gexec)
        return fail(YTRC_EINVAL, "ytrc_batch_request_ms: per-request times exist for plain, uncaptured batches only");
    HIPCHK(hipEventSynchronize(b->ev_req[2 * i + 1]));
    HIPCHK(hipEventElapsedTime(ms, b->ev_req[2 * i], b->ev_req[2 * i + 1]));
    return YTRC_OK;
}

namespace {

// The end of a batch's last launch: its GPU time, or YTRC_ECANCELLED when
// it was asked to stop.
int batch_finish(ytrc_batch *b, float *ms_out)
{
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (ms_out) *ms_out = ms;
    if (b->cancelled)
        return fail(YTRC_ECANCELLED, "ytrc_batch_sync: the launch was stopped (ytrc_batch_cancel); "
                                     "its images are incomplete");
    b->total_ms += ms;
    b->timed += 1;
    return YTRC_OK;
}

}  // namespace

int ytrc_batch_sync(ytrc_batch *b, float *ms_out)
{
    if (b == nullptr) return fail(YTRC_EINVAL, "ytrc_batch_sync: NULL batch");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipEventSynchronize(b->ev1));
    return batch_finish(b, ms_out);
}

int ytrc_batch_sync_for(ytrc_batch *b, double timeout_ms, float *ms_out)
{
    if (b == nullptr) return fail(YTRC_EINVAL, "ytrc_batch_sync_for: NULL batch");
    if (!(timeout_ms >= 0.0)) return fail(YTRC_EINVAL, "ytrc_batch_sync_for: timeout_ms must be >= 0");
    HIPCHK(hipSetDevice(b->ctx->device));
    const auto t0 = std::chrono::steady_clock::now();
    while (true) {
        const hipError_t e = hipEventQuery(b->ev1);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) return fail(YTRC_EDEVICE, std::string("hipEventQuery: ") + hipGetErrorString(e));
        const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (el >= timeout_ms)
            return fail(YTRC_ETIMEDOUT, "ytrc_batch_sync_for: the launch is still running after " +
                                            std::to_string(timeout_ms) + " ms");
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    return batch_finish(b, ms_out);
}

int ytrc_batch_cancel(ytrc_batch *b)
{
    if (b == nullptr) return fail(YTRC_EINVAL, "ytrc_batch_cancel: NULL batch");
    __atomic_store_n(b->h_stop, 1, __ATOMIC_SEQ_CST);
    b->cancelled = true;
    return YTRC_OK;
}

int ytrc_batch_kernel_ms(const ytrc_batch *b, float *avg_ms, int32_t *n_launches)
{
    if (b == nullptr) return fail(YTRC_EINVAL, "ytrc_batch_kernel_ms: NULL batch");
    if (avg_ms) *avg_ms = b->timed ? (float)(b->total_ms / b->timed) : 0.0f;
    if (n_launches) *n_launches = b->timed;
    return YTRC_OK;
}

int ytrc_batch_read_image(ytrc_batch *b, int32_t i, double *image_out)
{
    if (b == nullptr || image_out == nullptr) return fail(YTRC_EINVAL, "ytrc_batch_read_image: NULL");
    if (i < 0 || i >= (int)b->reqs.size()) return fail(YTRC_EINVAL, "ytrc_batch_read_image: bad index");
    if (b->cancelled)
        return fail(YTRC_ECANCELLED, "ytrc_batch_read_image: the last launch was stopped; launch again");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    return read_request_image(b->reqs[i], b->h_imgs[i], image_out);
}

int ytrc_batch_read_partial(ytrc_batch *b, int32_t i, double *map_out)
{
    if (b == nullptr || map_out == nullptr) return fail(YTRC_EINVAL, "ytrc_batch_read_partial: NULL");
    if (i < 0 || i >= (int)b->reqs.size()) return fail(YTRC_EINVAL, "ytrc_batch_read_partial: bad index");
    const ReqHost &q = b->reqs[i];
    if (!(q.r.flags & YTRC_FLAG_PARTIAL))
        return fail(YTRC_EINVAL, "ytrc_batch_read_partial: request was not built with YTRC_FLAG_PARTIAL");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    std::vector<float> h((size_t)q.npix * 8);
    HIPCHK(hipMemcpy(h.data(), b->h_imgs[i].out, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
    for (int64_t p = 0; p < q.npix; p++)
        for (int c = 0; c < 8; c++) map_out[(size_t)p * 8 + c] = (double)h[(size_t)c * q.npix + p];
    return YTRC_OK;
}

int ytrc_batch_read_record(ytrc_batch *b, int32_t i, int64_t *visits, uint64_t *hashes,
                           int64_t *samples)
{
    if (b == nullptr) return fail(YTRC_EINVAL, "ytrc_batch_read_record: NULL batch");
    if (i < 0 || i >= (int)b->reqs.size()) return fail(YTRC_EINVAL, "ytrc_batch_read_record: bad index");
    const DImg &d = b->h_imgs[i];
    if (d.rec_visits == nullptr)
        return fail(YTRC_EINVAL, "ytrc_batch_read_record: request was not built with YTRC_FLAG_RECORD");
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    const size_t n = (size_t)b->reqs[i].npix;
    if (visits) HIPCHK(hipMemcpy(visits, d.rec_visits, 8 * n, hipMemcpyDeviceToHost));
    if (hashes) HIPCHK(hipMemcpy(hashes, d.rec_hash, 8 * n, hipMemcpyDeviceToHost));
    if (samples) HIPCHK(hipMemcpy(samples, d.rec_samples, 8 * n, hipMemcpyDeviceToHost));
    return YTRC_OK;
}

int ytrc_hbm_probe(ytrc_ctx *ctx, uint64_t bytes, int32_t iters, float *gbps, float *ms_per_copy)
{
    if (ctx == nullptr || bytes < 16 || iters <= 0) return fail(YTRC_EINVAL, "ytrc_hbm_probe: bad argument");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)(bytes / 16);
    float4 *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, n * 16));
    hipError_t e = hipMalloc(&b, n * 16);
    if (e != hipSuccess) {
        hipFree(a);
        return fail(YTRC_EDEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const int grid = ctx->n_cu * 8;
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, ctx->stream, (const float4 *)a, b, n);
    HIPCHK(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < iters; i++)
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, ctx->stream, (const float4 *)a, b, n);
    HIPCHK(hipEventRecord(e1, ctx->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    const double per = ms / iters;
    if (ms_per_copy) *ms_per_copy = (float)per;
    if (gbps) *gbps = (float)(2.0 * (double)n * 16.0 / (per * 1e-3) / 1e9);
    return YTRC_OK;
}

int ytrc_batch_bytes(const ytrc_batch *b, uint64_t *bytes)
{
    if (b == nullptr || bytes == nullptr) return fail(YTRC_EINVAL, "ytrc_batch_bytes: NULL");
    *bytes = b->arena_bytes;
    return YTRC_OK;
}

void ytrc_batch_destroy(ytrc_batch *b)
{
    if (b == nullptr) return;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    if (b->arena) {
        hipFree(b->arena);
        std::lock_guard<std::mutex> lk(b->ctx->mu);
        b->ctx->used -= b->arena_bytes;
    }
    {
        std::lock_guard<std::mutex> lk(b->ctx->mu);
        for (int id : b->ds_ids) {
            auto it = b->ctx->datasets.find(id);
            if (it != b->ctx->datasets.end()) it->second.refs--;
        }
    }
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->ev_fork) hipEventDestroy(b->ev_fork);
    if (b->ev_join) hipEventDestroy(b->ev_join);
    for (auto &e : b->ev_req) hipEventDestroy(e);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->h_stop) hipHostFree(b->h_stop);
    delete b;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Serving: continuous batching of requests against resident datasets.
//
// Admitted requests get a descriptor slot and a region of a pre-allocated
// device arena (the HBM budget), and their 8x8 ray tiles join a FIFO.  The
// host keeps up to two chunks in flight; a chunk takes the next tiles of the
// FIFO -- from every admitted request, in arrival order -- and runs them as
// persistent tile-queue launches (one per kernel variant).  A request that
// arrives while others are on the GPU joins at the next chunk boundary; it
// is finished when the chunk holding its last tile completes.  Every kernel
// terminates; nothing spins on host-written flags.

namespace {

struct ArenaAlloc {
    // First-fit allocator over one device allocation (no hipMalloc/hipFree
    // while serving).
    char *base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    void init(char *b, size_t n)
    {
        base = b;
        size = n;
        free_.clear();
        free_[0] = n;
    }
    bool alloc(size_t n, size_t &off)
    {
        n = align_up(n, 256);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                const size_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = align_up(n, 256);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct SrvReq {
    int64_t id = 0;
    PrepReq P;
    DImg d;
    size_t off = 0;
    int slot = -1;
    int variant = -1;
    int next_tile = 0;
    int64_t last_chunk = -1;
    int state = 0;  // 0 tiles pending, 1 all tiles launched, 2 done, 3 cancelled
    int tf_entry = -1;
    bool ds_ref = false;  // holds a reference on its dataset
    std::chrono::steady_clock::time_point t_submit;
};

struct TFEntry {
    std::string key;
    int base = 0;
    int len = 0;
    int refs = 0;
};

struct SrvChunk {
    int64_t seq = -1;
    hipEvent_t done = nullptr;
    int2 *h_items = nullptr;  // pinned
    int2 *d_items = nullptr;
    unsigned int *d_counters = nullptr;
    std::vector<int64_t> finishing;
    bool busy = false;
};

constexpr int SRV_CHUNKS = 3;
constexpr int SRV_MAX_VARIANTS = 64;  // distinct persistent kernels one server launches

}  // namespace

struct ytrc_server {
    ytrc_ctx *ctx = nullptr;
    int max_slots = 0;
    int chunk_tiles = 0;
    int max_inflight = 2;
    int policy = YTRC_SERVE_FIFO;
    double age_limit_ms = 0.0;
    uint64_t budget = 0;
    char *arena = nullptr;
    ArenaAlloc alloc;
    DImg *d_slots = nullptr;
    std::vector<int> free_slots;
    float4 *d_tf = nullptr;
    int tf_cap = 0;  // float4 entries of the table: the LDS a workgroup may allocate
    std::vector<TFEntry> tfs;
    int tf_hi = 0;
    std::map<int64_t, SrvReq> reqs;
    std::deque<int64_t> fifo;
    std::vector<std::tuple<int, int, int, int>> vkeys;
    std::vector<pfn_t> vfns;
    SrvChunk chunks[SRV_CHUNKS];
    std::deque<int> inflight;
    int64_t next_seq = 0;
    int64_t next_id = 1;
    std::deque<int64_t> finished;
    uint64_t used = 0;
    int64_t tiles_launched = 0;
    int64_t chunks_launched = 0;
    std::mutex mu;
};

namespace {

int srv_variant(ytrc_server *S, const PrepReq &P)
{
    const bool rec = (P.q.r.flags & YTRC_FLAG_RECORD) != 0;
    auto key = std::make_tuple(P.q.r.kind, P.q.nfp, rec ? 1 : 0, P.q.mode);
    for (size_t i = 0; i < S->vkeys.size(); i++)
        if (S->vkeys[i] == key) return (int)i;
    if ((int)S->vkeys.size() >= SRV_MAX_VARIANTS) return -1;
    const pfn_t fn = pick_persist(P.q.r.kind, P.q.nfp, rec, P.q.mode);
    if (has_tf(P.q.r.kind) && S->tf_cap * sizeof(float4) > 64 * 1024) {
        // launches copy up to the whole table into LDS
        const hipError_t e = hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)(S->tf_cap * sizeof(float4)));
        if (e != hipSuccess) {
            fail(YTRC_EDEVICE, std::string("server: hipFuncSetAttribute: ") + hipGetErrorString(e));
            return -2;
        }
    }
    S->vkeys.push_back(key);
    S->vfns.push_back(fn);
    return (int)S->vkeys.size() - 1;
}

// Finds or places a transfer-function table in the server's device table.
int srv_tf_acquire(ytrc_server *S, const std::vector<float4> &t, int &entry)
{
    std::string key((const char *)t.data(), t.size() * sizeof(float4));
    for (size_t i = 0; i < S->tfs.size(); i++)
        if (S->tfs[i].len > 0 && S->tfs[i].key == key) {
            S->tfs[i].refs++;
            // launches copy the table's first tf_hi entries into LDS
            S->tf_hi = std::max(S->tf_hi, S->tfs[i].base + S->tfs[i].len);
            entry = (int)i;
            return YTRC_OK;
        }
    // first fit among gaps of the entries in use
    std::vector<std::pair<int, int>> used;
    for (auto &e : S->tfs)
        if (e.refs > 0) used.push_back({e.base, e.len});
    std::sort(used.begin(), used.end());
    int pos = 0;
    for (auto &u : used) {
        if (u.first - pos >= (int)t.size()) break;
        pos = std::max(pos, u.first + u.second);
    }
    if ((int)t.size() > S->tf_cap)
        return fail(YTRC_EINVAL, "server: the transfer-function table needs " +
                                     std::to_string(t.size() * sizeof(float4)) + " B of LDS, more than the " +
                                     std::to_string(S->tf_cap * sizeof(float4)) +
                                     " B a workgroup may allocate on this device");
    if (pos + (int)t.size() > S->tf_cap)
        return fail(YTRC_EBUDGET, "server: no room for another transfer function in LDS");
    // Released tables keep their place (and are found again by key) until a
    // new table is written over them: forget every released table the new
    // one overlaps.
    for (auto &e : S->tfs)
        if (e.refs == 0 && e.len > 0 && e.base < pos + (int)t.size() && pos < e.base + e.len) {
            e.len = 0;
            e.key.clear();
        }
    HIPCHK(hipMemcpy(S->d_tf + pos, t.data(), t.size() * sizeof(float4), hipMemcpyHostToDevice));
    // A free entry: one forgotten above, else any released one.
    int slot = -1;
    for (size_t i = 0; i < S->tfs.size() && slot < 0; i++)
        if (S->tfs[i].refs == 0 && S->tfs[i].len == 0) slot = (int)i;
    for (size_t i = 0; i < S->tfs.size() && slot < 0; i++)
        if (S->tfs[i].refs == 0) slot = (int)i;
    if (slot < 0) {
        S->tfs.push_back(TFEntry());
        slot = (int)S->tfs.size() - 1;
    }
    S->tfs[slot].key = key;
    S->tfs[slot].base = pos;
    S->tfs[slot].len = (int)t.size();
    S->tfs[slot].refs = 1;
    S->tf_hi = 0;
    for (auto &e : S->tfs)
        if (e.refs > 0) S->tf_hi = std::max(S->tf_hi, e.base + e.len);
    entry = slot;
    return YTRC_OK;
}

void srv_release(ytrc_server *S, SrvReq &R)
{
    if (R.off != (size_t)-1 && R.P.bytes) {
        S->alloc.release(R.off, R.P.bytes);
        S->used -= align_up(R.P.bytes, 256);
        R.off = (size_t)-1;
    }
    if (R.slot >= 0) {
        S->free_slots.push_back(R.slot);
        R.slot = -1;
    }
    if (R.tf_entry >= 0) {
        S->tfs[R.tf_entry].refs--;
        R.tf_entry = -1;
    }
    if (R.ds_ref) {
        std::lock_guard<std::mutex> lc(S->ctx->mu);
        auto it = S->ctx->datasets.find(R.P.q.r.ds_id);
        if (it != S->ctx->datasets.end()) it->second.refs--;
        R.ds_ref = false;
    }
}

// Estimated work left in a request, in projection-tile units: a tile of a
// sampler that interpolates (volume render, interpolated projection,
// light source) costs about 4 projection tiles on the 512^3 benchmark
// (54 vs 12 ms for the same 32 cameras).
double srv_work_left(const SrvReq &R)
{
    const double w = R.P.q.r.kind == YTRC_PROJECTION ? 1.0 : 4.0;
    return w * (double)(R.P.q.ntiles - R.next_tile);
}

// The order in which requests with tiles pending fill the next chunk.
std::vector<int64_t> srv_fill_order(ytrc_server *S)
{
    std::vector<int64_t> ids(S->fifo.begin(), S->fifo.end());
    if (S->policy != YTRC_SERVE_SHORTEST_FIRST) return ids;
    const auto now = std::chrono::steady_clock::now();
    struct Key {
        bool aged;
        double work;
        size_t pos;  // arrival rank
        int64_t id;
    };
    std::vector<Key> keys;
    keys.reserve(ids.size());
    for (size_t i = 0; i < ids.size(); i++) {
        const SrvReq &R = S->reqs.at(ids[i]);
        const double age = std::chrono::duration<double, std::milli>(now - R.t_submit).count();
        keys.push_back(Key{age > S->age_limit_ms, srv_work_left(R), i, ids[i]});
    }
    // Aged requests first, in arrival order; then the least work left.
    std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
        if (a.aged != b.aged) return a.aged;
        if (!a.aged && a.work != b.work) return a.work < b.work;
        return a.pos < b.pos;
    });
    for (size_t i = 0; i < keys.size(); i++) ids[i] = keys[i].id;
    return ids;
}

// Launches chunks while fewer than max_inflight are on the GPU.
int srv_pump(ytrc_server *S)
{
    hipStream_t s0 = S->ctx->stream;
    while ((int)S->inflight.size() < S->max_inflight && !S->fifo.empty()) {
        int ci = -1;
        for (int i = 0; i < SRV_CHUNKS; i++)
            if (!S->chunks[i].busy) {
                ci = i;
                break;
            }
        if (ci < 0) break;
        SrvChunk &C = S->chunks[ci];
        // Consecutive chunks alternate between the context's two streams, so
        // that a chunk's launches start in the tail of the one before it.
        // Chunks still retire in launch order (srv_poll), so a request is
        // finished only once every chunk that holds its tiles has retired.
        hipStream_t s = (S->chunks_launched & 1) ? S->ctx->stream2 : s0;
        C.finishing.clear();
        const int nv = (int)S->vkeys.size();
        std::vector<std::vector<int2>> per(nv);
        int filled = 0;
        bool launched_all = false;
        for (int64_t id : srv_fill_order(S)) {
            if (filled >= S->chunk_tiles) break;
            SrvReq &R = S->reqs.at(id);
            const int take = std::min(R.P.q.ntiles - R.next_tile, S->chunk_tiles - filled);
            for (int t = 0; t < take; t++) per[R.variant].push_back(make_int2(R.slot, R.next_tile + t));
            R.next_tile += take;
            filled += take;
            R.last_chunk = S->next_seq;
            if (R.next_tile >= R.P.q.ntiles) {
                R.state = 1;
                C.finishing.push_back(R.id);
                launched_all = true;
            }
        }
        if (launched_all)
            S->fifo.erase(std::remove_if(S->fifo.begin(), S->fifo.end(),
                                         [&](int64_t id) { return S->reqs.at(id).state == 1; }),
                          S->fifo.end());
        std::vector<int> voff(nv, 0), vcnt(nv, 0);
        int n = 0;
        for (int v = 0; v < nv; v++) {
            voff[v] = n;
            vcnt[v] = (int)per[v].size();
            for (auto &it : per[v]) C.h_items[n++] = it;
        }
        HIPCHK(hipMemcpyAsync(C.d_items, C.h_items, sizeof(int2) * n, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_zero_counters, dim3((SRV_MAX_VARIANTS * QCOUNTERS + 255) / 256), dim3(256), 0, s,
                           C.d_counters, SRV_MAX_VARIANTS * QCOUNTERS);
        HIPCHK(hipGetLastError());
        for (int v = 0; v < nv; v++) {
            if (vcnt[v] == 0) continue;
            const bool vr = has_tf(std::get<0>(S->vkeys[v]));
            const int ww = required_wave_words(std::get<0>(S->vkeys[v]), std::get<1>(S->vkeys[v]),
                                               std::get<2>(S->vkeys[v]) != 0, std::get<3>(S->vkeys[v]));
            const size_t lds = (vr ? sizeof(float4) * S->tf_hi : 0) + 4 * (size_t)ww * sizeof(unsigned);
            int nb = 0;
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, S->vfns[v], 256, lds));
            const int grid = std::max(1, std::min(std::max(1, nb) * S->ctx->n_cu, (vcnt[v] + 3) / 4));
            hipLaunchKernelGGL(S->vfns[v], dim3(grid), dim3(256), lds, s, (const DImg *)S->d_slots,
                               (const int2 *)(C.d_items + voff[v]), vcnt[v], C.d_counters + (size_t)v * QCOUNTERS,
                               (const float4 *)S->d_tf, vr ? S->tf_hi : 0, ww);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipEventRecord(C.done, s));
        C.seq = S->next_seq++;
        C.busy = true;
        S->inflight.push_back(ci);
        S->tiles_launched += n;
        S->chunks_launched += 1;
    }
    return YTRC_OK;
}

// Retires completed chunks; requests whose last tile ran in them finish.
int srv_poll(ytrc_server *S, bool block)
{
    while (!S->inflight.empty()) {
        SrvChunk &C = S->chunks[S->inflight.front()];
        hipError_t e = block ? hipEventSynchronize(C.done) : hipEventQuery(C.done);
        if (e == hipErrorNotReady) break;
        if (e != hipSuccess) return fail(YTRC_EDEVICE, std::string("server chunk: ") + hipGetErrorString(e));
        for (int64_t id : C.finishing) {
            auto it = S->reqs.find(id);
            if (it == S->reqs.end()) continue;
            if (it->second.state == 3) {
                srv_release(S, it->second);
                S->reqs.erase(it);
            } else {
                it->second.state = 2;
                S->finished.push_back(id);
            }
        }
        C.busy = false;
        S->inflight.pop_front();
        if (block) break;
    }
    return YTRC_OK;
}

}  // namespace

extern "C" {

int ytrc_server_create(ytrc_ctx *ctx, int32_t max_requests, uint64_t budget_bytes,
                       int32_t chunk_tiles, ytrc_server **out)
{
    if (ctx == nullptr || out == nullptr) return fail(YTRC_EINVAL, "ytrc_server_create: NULL argument");
    if (max_requests <= 0 || max_requests > 4096)
        return fail(YTRC_EINVAL, "ytrc_server_create: max_requests must be in 1..4096");
    if (chunk_tiles < 64 || chunk_tiles > (1 << 22))
        return fail(YTRC_EINVAL, "ytrc_server_create: chunk_tiles must be in 64..4194304");
    if (budget_bytes < (1u << 20)) return fail(YTRC_EINVAL, "ytrc_server_create: budget below 1 MiB");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    if (ctx->budget != 0 && ctx->used + budget_bytes > ctx->budget)
        return fail(YTRC_EBUDGET, "ytrc_server_create: server budget exceeds the context budget");
    ytrc_server *S = new ytrc_server();
    S->ctx = ctx;
    S->max_slots = max_requests;
    S->chunk_tiles = chunk_tiles;
    S->budget = budget_bytes;
    auto bad = [&](hipError_t e, const char *what) {
        if (S->arena) hipFree(S->arena);
        if (S->d_slots) hipFree(S->d_slots);
        if (S->d_tf) hipFree(S->d_tf);
        for (auto &c : S->chunks) {
            if (c.h_items) hipHostFree(c.h_items);
            if (c.d_items) hipFree(c.d_items);
            if (c.d_counters) hipFree(c.d_counters);
            if (c.done) hipEventDestroy(c.done);
        }
        delete S;
        return fail(YTRC_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e = hipMalloc(&S->arena, budget_bytes);
    if (e != hipSuccess) return bad(e, "hipMalloc(arena)");
    S->alloc.init(S->arena, budget_bytes);
    e = hipMalloc(&S->d_slots, sizeof(DImg) * max_requests);
    if (e != hipSuccess) return bad(e, "hipMalloc(slots)");
    S->tf_cap = (int)(ctx->max_lds / sizeof(float4));
    e = hipMalloc(&S->d_tf, sizeof(float4) * S->tf_cap);
    if (e != hipSuccess) return bad(e, "hipMalloc(tf)");
    for (auto &c : S->chunks) {
        e = hipHostMalloc(&c.h_items, sizeof(int2) * chunk_tiles, hipHostMallocDefault);
        if (e != hipSuccess) return bad(e, "hipHostMalloc(items)");
        e = hipMalloc(&c.d_items, sizeof(int2) * chunk_tiles);
        if (e != hipSuccess) return bad(e, "hipMalloc(items)");
        e = hipMalloc(&c.d_counters, sizeof(unsigned int) * SRV_MAX_VARIANTS * QCOUNTERS);
        if (e != hipSuccess) return bad(e, "hipMalloc(counters)");
        e = hipEventCreateWithFlags(&c.done, hipEventDisableTiming);
        if (e != hipSuccess) return bad(e, "hipEventCreate");
    }
    for (int i = max_requests - 1; i >= 0; i--) S->free_slots.push_back(i);
    ctx->used += budget_bytes;
    *out = S;
    return YTRC_OK;
}

int ytrc_server_submit(ytrc_server *S, const ytrc_request *r, int64_t *id_out)
{
    if (S == nullptr || r == nullptr || id_out == nullptr) return fail(YTRC_EINVAL, "ytrc_server_submit: NULL");
    if (r->flags & (YTRC_FLAG_RECORD | YTRC_FLAG_PARTIAL))
        return fail(YTRC_EINVAL, "ytrc_server_submit: record mode and partial maps are for batches");
    std::lock_guard<std::mutex> lk(S->mu);
    HIPCHK(hipSetDevice(S->ctx->device));
    SrvReq R;
    {
        std::lock_guard<std::mutex> lc(S->ctx->mu);
        const int rc = prep_request(S->ctx, *r, "request: ", R.P);
        if (rc != YTRC_OK) return rc;
        // Held until the request is read or cancelled (srv_release).
        S->ctx->datasets[r->ds_id].refs++;
        R.ds_ref = true;
    }
    // Admission: a descriptor slot and arena room for the request's region.
    R.off = (size_t)-1;
    if (S->free_slots.empty()) {
        srv_release(S, R);
        return fail(YTRC_EBUDGET, "server: all request slots are in use");
    }
    if (align_up(R.P.bytes, 256) > S->budget) {
        srv_release(S, R);
        return fail(YTRC_EINVAL, "server: request needs " + std::to_string(R.P.bytes) +
                                     " bytes, more than the server's whole budget of " +
                                     std::to_string(S->budget));
    }
    size_t off = 0;
    if (!S->alloc.alloc(R.P.bytes, off)) {
        srv_release(S, R);
        return fail(YTRC_EBUDGET, "server: request needs " + std::to_string(R.P.bytes) +
                                      " bytes, more than the free HBM budget");
    }
    R.off = off;
    S->used += align_up(R.P.bytes, 256);
    int tf_base = 0;
    if (has_tf(r->kind)) {
        const int rc = srv_tf_acquire(S, R.P.tft, R.tf_entry);
        if (rc != YTRC_OK) {
            srv_release(S, R);
            return rc;
        }
        tf_base = S->tfs[R.tf_entry].base;
    }
    R.slot = S->free_slots.back();
    S->free_slots.pop_back();
    std::vector<float> stage;
    int rc = stage_request(S->ctx, R.P, S->arena + off, S->d_tf, tf_base, R.d, stage);
    if (rc == YTRC_OK) {
        hipError_t e = hipMemcpy(S->d_slots + R.slot, &R.d, sizeof(DImg), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(YTRC_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    if (rc != YTRC_OK) {
        srv_release(S, R);
        return rc;
    }
    R.variant = srv_variant(S, R.P);
    if (R.variant == -2) {
        srv_release(S, R);
        return YTRC_EDEVICE;
    }
    if (R.variant < 0) {
        srv_release(S, R);
        return fail(YTRC_ENOTIMPL, "server: too many distinct kernel variants");
    }
    R.id = S->next_id++;
    R.t_submit = std::chrono::steady_clock::now();
    R.P.q.r.order = nullptr;  // caller-owned pointers are not kept
    R.P.q.r.zbuffer = nullptr;
    R.P.q.r.image_in = nullptr;
    R.P.q.r.tf = nullptr;
    R.P.q.r.pix_pos = nullptr;
    R.P.q.r.pix_dir = nullptr;
    R.P.q.r.camera_data = nullptr;
    *id_out = R.id;
    const int64_t id = R.id;
    S->reqs.emplace(id, std::move(R));
    S->fifo.push_back(id);
    return srv_pump(S);
}

int ytrc_server_progress(ytrc_server *S, int64_t *done_ids, int32_t cap, int32_t *n_done)
{
    if (S == nullptr || n_done == nullptr) return fail(YTRC_EINVAL, "ytrc_server_progress: NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    HIPCHK(hipSetDevice(S->ctx->device));
    int rc = srv_poll(S, false);
    if (rc != YTRC_OK) return rc;
    rc = srv_pump(S);
    if (rc != YTRC_OK) return rc;
    int n = 0;
    while (!S->finished.empty() && n < cap) {
        if (done_ids) done_ids[n] = S->finished.front();
        S->finished.pop_front();
        n++;
    }
    *n_done = n;
    return YTRC_OK;
}

int ytrc_server_wait(ytrc_server *S, int64_t *done_ids, int32_t cap, int32_t *n_done)
{
    if (S == nullptr || n_done == nullptr) return fail(YTRC_EINVAL, "ytrc_server_wait: NULL");
    {
        std::lock_guard<std::mutex> lk(S->mu);
        HIPCHK(hipSetDevice(S->ctx->device));
        if (S->finished.empty() && !S->inflight.empty()) {
            const int rc = srv_poll(S, true);
            if (rc != YTRC_OK) return rc;
        }
    }
    return ytrc_server_progress(S, done_ids, cap, n_done);
}

int ytrc_server_wait_for(ytrc_server *S, double timeout_ms, int64_t *done_ids, int32_t cap, int32_t *n_done)
{
    if (S == nullptr || n_done == nullptr) return fail(YTRC_EINVAL, "ytrc_server_wait_for: NULL");
    if (!(timeout_ms >= 0.0)) return fail(YTRC_EINVAL, "ytrc_server_wait_for: timeout_ms must be >= 0");
    const auto t0 = std::chrono::steady_clock::now();
    while (true) {
        const int rc = ytrc_server_progress(S, done_ids, cap, n_done);
        if (rc != YTRC_OK || *n_done > 0) return rc;
        bool idle;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            idle = S->inflight.empty() && S->fifo.empty();
        }
        if (idle) return YTRC_OK;  // nothing left to wait for
        const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (el >= timeout_ms)
            return fail(YTRC_ETIMEDOUT, "ytrc_server_wait_for: no request finished within " +
                                            std::to_string(timeout_ms) + " ms");
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

int ytrc_server_read(ytrc_server *S, int64_t id, double *image_out)
{
    if (S == nullptr || image_out == nullptr) return fail(YTRC_EINVAL, "ytrc_server_read: NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(YTRC_EINVAL, "ytrc_server_read: unknown request");
    if (it->second.state != 2) return fail(YTRC_EINVAL, "ytrc_server_read: request is not finished");
    HIPCHK(hipSetDevice(S->ctx->device));
    const int rc = read_request_image(it->second.P.q, it->second.d, image_out);
    srv_release(S, it->second);
    S->reqs.erase(it);
    return rc;
}

int ytrc_server_read_planar(ytrc_server *S, int64_t id, float *planes, int32_t cap_ch, int32_t *nch)
{
    if (S == nullptr || planes == nullptr || nch == nullptr)
        return fail(YTRC_EINVAL, "ytrc_server_read_planar: NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(YTRC_EINVAL, "ytrc_server_read_planar: unknown request");
    if (it->second.state != 2) return fail(YTRC_EINVAL, "ytrc_server_read_planar: request is not finished");
    const ReqHost &q = it->second.P.q;
    if (cap_ch < q.nch_out) return fail(YTRC_EINVAL, "ytrc_server_read_planar: output has too few channels");
    HIPCHK(hipSetDevice(S->ctx->device));
    HIPCHK(hipMemcpy(planes, it->second.d.out, sizeof(float) * (size_t)q.npix * q.nch_out,
                     hipMemcpyDeviceToHost));
    if (!q.image_in.empty())
        for (int c = 0; c < q.nch_out; c++)
            for (int64_t p = 0; p < q.npix; p++) {
                float &o = planes[(size_t)c * q.npix + p];
                o = (float)merge_input(q.r.kind, o, q.image_in[(size_t)p * 4 + c]);
            }
    *nch = q.nch_out;
    srv_release(S, it->second);
    S->reqs.erase(it);
    return YTRC_OK;
}

int ytrc_server_set_policy(ytrc_server *S, int32_t policy, double age_limit_ms)
{
    if (S == nullptr) return fail(YTRC_EINVAL, "ytrc_server_set_policy: NULL");
    if (policy != YTRC_SERVE_FIFO && policy != YTRC_SERVE_SHORTEST_FIRST)
        return fail(YTRC_EINVAL, "ytrc_server_set_policy: unknown policy " + std::to_string(policy));
    if (!(age_limit_ms >= 0.0))
        return fail(YTRC_EINVAL, "ytrc_server_set_policy: age_limit_ms must be >= 0");
    std::lock_guard<std::mutex> lk(S->mu);
    S->policy = policy;
    S->age_limit_ms = age_limit_ms;
    return YTRC_OK;
}

int ytrc_server_cancel(ytrc_server *S, int64_t id)
{
    if (S == nullptr) return fail(YTRC_EINVAL, "ytrc_server_cancel: NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(YTRC_EINVAL, "ytrc_server_cancel: unknown request");
    SrvReq &R = it->second;
    if (R.state == 0) {
        S->fifo.erase(std::remove(S->fifo.begin(), S->fifo.end(), id), S->fifo.end());
        // Tiles already launched may still be running: release with the
        // chunk that holds its last launched tile, or now if none ran.
        bool running = false;
        for (int ci : S->inflight)
            if (S->chunks[ci].seq == R.last_chunk) {
                S->chunks[ci].finishing.push_back(id);
                running = true;
            }
        if (!running) {
            srv_release(S, R);
            S->reqs.erase(it);
            return YTRC_OK;
        }
        R.state = 3;
    } else if (R.state == 1) {
        R.state = 3;  // released when its last chunk retires
    } else if (R.state == 2) {
        S->finished.erase(std::remove(S->finished.begin(), S->finished.end(), id), S->finished.end());
        srv_release(S, R);
        S->reqs.erase(it);
    }
    return YTRC_OK;
}

int ytrc_server_stats(const ytrc_server *S, int64_t *stats)
{
    if (S == nullptr || stats == nullptr) return fail(YTRC_EINVAL, "ytrc_server_stats: NULL");
    stats[0] = (int64_t)S->reqs.size();
    stats[1] = (int64_t)S->fifo.size();
    stats[2] = (int64_t)S->inflight.size();
    stats[3] = (int64_t)S->used;
    stats[4] = (int64_t)S->budget;
    stats[5] = S->tiles_launched;
    stats[6] = S->chunks_launched;
    stats[7] = (int64_t)S->free_slots.size();
    return YTRC_OK;
}

void ytrc_server_destroy(ytrc_server *S)
{
    if (S == nullptr) return;
    hipSetDevice(S->ctx->device);
    hipStreamSynchronize(S->ctx->stream);
    hipStreamSynchronize(S->ctx->stream2);
    hipFree(S->arena);
    hipFree(S->d_slots);
    hipFree(S->d_tf);
    for (auto &c : S->chunks) {
        hipHostFree(c.h_items);
        hipFree(c.d_items);
        hipFree(c.d_counters);
        hipEventDestroy(c.done);
    }
    {
        std::lock_guard<std::mutex> lk(S->ctx->mu);
        S->ctx->used -= S->budget;
        // Requests never read or cancelled give their datasets back.
        for (auto &kv : S->reqs) {
            if (!kv.second.ds_ref) continue;
            auto it = S->ctx->datasets.find(kv.second.P.q.r.ds_id);
            if (it != S->ctx->datasets.end()) it->second.refs--;
        }
    }
    delete S;
}

}  // extern "C"
