// ytrc_device.h -- device-side data layout and ray-casting core for gfx950.
//
// Compiled with -ffp-contract=off: every FP64 operation of the DDA and of
// the ray set-up is a separately rounded IEEE operation, in the reference's
// order, so cell-visit sequences are bit-identical to yt's walk_volume
// (grid_traversal.pyx:26-140).  FP32 sampling code uses explicit fmaf.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ytrc {

// Brick descriptor in HBM (VolumeContainer, volume_container.pxd:13-33).
// 128 bytes; read with scalar loads because every lane of a wave walks the
// same brick at the same time.
struct alignas(16) DBrick {
    double le[3];
    double re[3];
    double dds[3];
    double idds[3];
    int32_t dims[3];
    int32_t vertex;      // 1: (dims+1)^3 points per field
    int64_t data_off;    // float offset of the brick's points in the pool
    int64_t mask_off;    // byte offset of the dims^3 mask, -1 = none (all ones)
};

// Transfer function parameters (FieldInterpolationTable x n_fits,
// field_interpolation_tables.pxd:27-39).  Tables live in LDS as float4
// {y, dy, d0, 0} per (table, bin).
struct alignas(16) DTF {
    int32_t n_fits;
    int32_t grey;
    int32_t total_bins;
    int32_t mode;         // 1: ColorTransferFunction layout, 0: general
    int32_t nbins[6];
    int32_t off[6];
    int32_t field_id[6];
    int32_t wfield[6];
    int32_t wtable[6];
    int32_t ftable[6];
    int32_t chexp[4];     // per channel: 2-bit power of each raw table value, -1 = zero
    float b0[6];
    float b1[6];
    float idbin[6];
    // ColorTransferFunction tables in transmittance form: bin-coordinate
    // offset (u = v * idbin + qc1) and the index of the all-zero entry.
    float qc1, qzb;
    const float4 *table;  // total_bins float4 in global memory
};

// One image request on the device.
struct alignas(16) DImg {
    double vp_pos[12];
    double vp_dir[3];
    double width[3];
    const DBrick *bricks;
    const float *pool;
    const uint8_t *masks;
    const int32_t *order;   // n_order brick indices
    const int4 *ext;        // n_order clipped extents {i0, i1, j0, j1}
    const double *zbuf;     // nv0*nv1 or null
    const float *init;      // 4 planes of nv0*nv1 or null
    float *out;             // 4 planes of nv0*nv1
    int64_t *rec_visits;    // record mode only
    uint64_t *rec_hash;
    int64_t *rec_samples;
    const DTF *tf;
    double ldir[3];         // light-source sampler: unit light direction
    double lrgba[4];        //   and light colour
    const double *pix_pos;  // per-pixel lens: nv0*nv1*3 origins, else null
    const double *pix_dir;  // per-pixel lens: nv0*nv1*3 directions
    const int *stop;        // host-pinned stop flag of the launch, or null
    int32_t nv0, nv1;
    int32_t kind, n_samples;
    int32_t n_order, nf;
    int32_t nfp, tiles0;
    int32_t tiles1, flags;
    int32_t nch_out, tf_slot;
    int32_t tform;          // 1: VR compositing state kept as transmittance
    int32_t tsh;            // tile = (64 >> tsh) rows x (1 << tsh) columns of rays
    double inv_ns;
};

// Cell-centred bricks (the projection path) are stored in blocks of 4x4x4
// cells: blocks in C order over the brick padded to whole blocks, cells in C
// order inside a block (256 B per block and field).  A ray bundle walking
// through a block touches two 128-B lines instead of up to 16 rows of the C
// order.
__host__ __device__ inline int64_t blocked_offset(int i, int j, int k, int nbj, int nbk)
{
    const int64_t blk = ((int64_t)(i >> 2) * nbj + (j >> 2)) * nbk + (k >> 2);
    return (blk << 6) | ((i & 3) << 4) | ((j & 3) << 2) | (k & 3);
}

// Vertex-centred bricks (the volume-render, interpolated-projection and
// light-source paths) are stored in blocks of VB x VB x VB cells.  Each
// block holds the (VB+1)^3 vertices its cells read -- vertices on a face
// shared by two blocks are stored in both -- padded to VBP points; blocks in
// C order over the brick padded to whole blocks, vertices in C order inside
// a block.  A cell's 8 corners then lie in one block at fixed offsets
// (vcorner), so a visit addresses them from one base, and a ray crossing a
// block reads its four 128-B lines instead of two new C-order rows per step
// across the rows.  Blocks of 4^3 cells (125 vertices in 512 B) store 1.95x
// the vertices; against the C order they cut the volume render's time at
// 512^3 by 20 % (by half where rays are sparser than cells), and beat 2^3
// blocks (27 vertices in 128 B).
constexpr int VB_SHIFT = 2;                 // log2 of the cells per block edge
constexpr int VB = 1 << VB_SHIFT;
constexpr int VB1 = VB + 1;                 // vertices per block edge
constexpr int VBP_SHIFT = VB == 2 ? 5 : 7;  // 27 -> 32, 125 -> 128 points
constexpr int VBP = 1 << VBP_SHIFT;
static_assert(VB1 * VB1 * VB1 <= VBP, "vertex block padding");

__host__ __device__ inline int vblocks(int cells) { return (cells + VB - 1) >> VB_SHIFT; }

// Point offset of corner (0, 0, 0) of cell (i, j, k).
__host__ __device__ inline int64_t vertex_block_offset(int i, int j, int k, int nbj, int nbk)
{
    const int64_t blk = ((int64_t)(i >> VB_SHIFT) * nbj + (j >> VB_SHIFT)) * nbk + (k >> VB_SHIFT);
    return (blk << VBP_SHIFT) + ((i & (VB - 1)) * VB1 + (j & (VB - 1))) * VB1 + (k & (VB - 1));
}

// Offset of the block's vertex (a, b, c), 0 <= a, b, c <= VB, from its
// vertex (0, 0, 0); for a, b, c in {0, 1}, the cell corner offsets.
__host__ __device__ constexpr int vcorner(int a, int b, int c) { return (a * VB1 + b) * VB1 + c; }

inline int64_t stored_points(const DBrick &h)
{
    if (h.vertex)
        return (int64_t)vblocks(h.dims[0]) * vblocks(h.dims[1]) * vblocks(h.dims[2]) * VBP;
    return (int64_t)((h.dims[0] + 3) >> 2) * ((h.dims[1] + 3) >> 2) * ((h.dims[2] + 3) >> 2) * 64;
}

// One 64-lane wavefront casts one tile of 64 rays: 8x8 by default, 4x16 or
// 16x4 when a batch is built with another shape.
constexpr int TILE_SHIFT_DEFAULT = 3;

__device__ __forceinline__ double y_fmin(double a, double b) { return (a < b) ? a : b; }

// Marks a value the compiler cannot prove uniform as wave-uniform (scalar
// register); only for values that are equal in every lane.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uni_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double y_fmax(double a, double b) { return (a > b) ? a : b; }

// The host's stop request for a launch (ytrc_batch_cancel), read from
// host-pinned memory at system scope by persistent waves (the analog of the
// reference's PyErr_CheckSignals every 1000 rays, image_samplers.pyx:188-190).
__device__ __forceinline__ bool stop_requested(const int *stop)
{
    return stop != nullptr && __hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0;
}

// generate_vector_info_plane_parallel: lenses.pyx:179-192.
__device__ __forceinline__ void ray_setup(const DImg &im, int vi, int vj, double pos[3],
                                          double dir[3])
{
    double px = im.width[0] * (double)vi / ((double)im.nv0 - 1.0) - im.width[0] / 2.0;
    double py = im.width[1] * (double)vj / ((double)im.nv1 - 1.0) - im.width[1] / 2.0;
    pos[0] = im.vp_pos[0] * px + im.vp_pos[3] * py + im.vp_pos[9];
    pos[1] = im.vp_pos[1] * px + im.vp_pos[4] * py + im.vp_pos[10];
    pos[2] = im.vp_pos[2] * px + im.vp_pos[5] * py + im.vp_pos[11];
    dir[0] = im.vp_dir[0];
    dir[1] = im.vp_dir[1];
    dir[2] = im.vp_dir[2];
}

// DDA state after walk_volume's initialisation (grid_traversal.pyx:42-118).
struct Walk {
    double tmax[3];
    double tdelta[3];
    int cur[3];
    int step[3];
    double enter_t;
};

// grid_traversal.pyx:42-118.  Returns false where walk_volume returns 0
// before its loop.  iv[i] = 1/v_dir[i] (0 where v_dir[i] == 0).
template <bool UNIFORM_DIR = true>
__device__ __forceinline__ bool walk_init(const DBrick &b, const double pos[3],
                                          const double dir[3], const double iv[3],
                                          double max_t, Walk &w)
{
    double intersect_t = 1.1;
    int direction = -1;
    if (b.le[0] <= pos[0] && pos[0] < b.re[0] && b.le[1] <= pos[1] && pos[1] < b.re[1] &&
        b.le[2] <= pos[2] && pos[2] < b.re[2]) {
        intersect_t = 0.0;
        direction = 3;
    }
    // Plane-parallel rays share v_dir, so step and tdelta are wave-uniform;
    // per-pixel lenses give every lane its own.
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int st = (dir[i] < 0) ? -1 : ((dir[i] == 0.0) ? 0 : 1);
        w.step[i] = UNIFORM_DIR ? uni(st) : st;
    }
    if (direction != 3) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (w.step[i] == 0) continue;
            const int x = (i + 1) % 3;
            const int y = (i + 2) % 3;
            double tl = ((w.step[i] > 0) ? (b.le[i] - pos[i]) : (b.re[i] - pos[i])) * iv[i];
            double tx = (pos[x] + tl * dir[x]);
            double ty = (pos[y] + tl * dir[y]);
            if (fabs(tx - b.le[x]) < 1e-10 * b.dds[x])
                tx = b.le[x];
            else if (fabs(tx - b.re[x]) < 1e-10 * b.dds[x])
                tx = b.re[x];
            if (fabs(ty - b.le[y]) < 1e-10 * b.dds[y])
                ty = b.le[y];
            else if (fabs(ty - b.re[y]) < 1e-10 * b.dds[y])
                ty = b.re[y];
            if (b.le[x] <= tx && tx <= b.re[x] && b.le[y] <= ty && ty <= b.re[y] && 0.0 <= tl &&
                tl < intersect_t) {
                direction = i;
                intersect_t = tl;
            }
        }
    }
    if (!((0.0 <= intersect_t) && (intersect_t < max_t))) return false;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double td = (double)w.step[i] * iv[i] * b.dds[i];
        w.tdelta[i] = UNIFORM_DIR ? uni_d(td) : td;
        double ty = 0.0;
        if (i == direction && w.step[i] > 0) {
            w.cur[i] = 0;
        } else if (i == direction && w.step[i] < 0) {
            w.cur[i] = b.dims[i] - 1;
        } else {
            double tx = intersect_t * dir[i] + pos[i];
            ty = ((tx - b.le[i]) * b.idds[i]);
            if (-1 < ty && ty < 0 && w.step[i] > 0)
                ty = 0.0;
            else if ((double)(b.dims[i] - 1) < ty && ty < (double)b.dims[i] && w.step[i] < 0)
                ty = (double)(b.dims[i] - 1);
            w.cur[i] = (int)(floor(ty));
        }
        if (w.step[i] > 0)
            ty = (double)(w.cur[i] + 1) * b.dds[i] + b.le[i];
        else if (w.step[i] < 0)
            ty = (double)w.cur[i] * b.dds[i] + b.le[i];
        w.tmax[i] = (w.step[i] == 0) ? 1e60 : (ty - pos[i]) * iv[i];
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (w.cur[i] == b.dims[i] && w.step[i] >= 0) return false;
        if (w.cur[i] == -1 && w.step[i] <= -1) return false;
    }
    w.enter_t = intersect_t;
    return true;
}

// The traversal loop, grid_traversal.pyx:120-138.  `sample` is called with
// (enter_t, exit_t, i, j, k) once per visited cell, in order.
template <class S>
__device__ __forceinline__ int walk_loop(const DBrick &b, Walk &w, double max_t, S &sample)
{
    int hit = 0;
    double t0 = w.tmax[0], t1 = w.tmax[1], t2 = w.tmax[2];
    int c0 = w.cur[0], c1 = w.cur[1], c2 = w.cur[2];
    double enter_t = w.enter_t;
    // Brick dims in scalar registers: the axis-indexed bound below would
    // otherwise be a memory load on every step.
    const int D0 = uni(b.dims[0]), D1 = uni(b.dims[1]), D2 = uni(b.dims[2]);
    while (true) {
        hit += 1;
        // The axis of the smallest tmax, ties to the later axis as the
        // reference's comparisons order it (grid_traversal.pyx:120-130):
        // tmax[0] < tmax[1] ? (tmax[0] < tmax[2] ? 0 : 2) : (tmax[1] <
        // tmax[2] ? 1 : 2), from two comparisons (m = the first pair's
        // choice, then m against tmax[2]) as lane masks (no axis index).
        const bool a01 = t0 < t1;
        const double m = a01 ? t0 : t1;
        const bool bm = m < t2;
        const bool s0 = a01 & bm;
        const bool s1 = !a01 & bm;
        const bool s2 = !(s0 | s1);
        const double tm = bm ? m : t2;
        // exit_t = fmin(tm, max_t) and the loop's enter_t >= max_t test on
        // the next enter_t (= tm), from one comparison
        const bool ge = tm >= max_t;
        const double exit_t = ge ? max_t : tm;
        sample(enter_t, exit_t, c0, c1, c2);
        // Step the chosen axis under an execution mask (the empty asm keeps
        // each update a masked region rather than selects) and test only
        // its new index.
        bool out = false;
        if (s0) {
            asm volatile("");
            c0 += w.step[0];
            t0 += w.tdelta[0];
            out = (unsigned)c0 >= (unsigned)D0;
        }
        if (s1) {
            asm volatile("");
            c1 += w.step[1];
            t1 += w.tdelta[1];
            out = (unsigned)c1 >= (unsigned)D1;
        }
        if (s2) {
            asm volatile("");
            c2 += w.step[2];
            t2 += w.tdelta[2];
            out = (unsigned)c2 >= (unsigned)D2;
        }
        enter_t = tm;
        if (out || ge) break;
    }
    return hit;
}

__device__ __forceinline__ bool in_brick(const DBrick &b, int c0, int c1, int c2)
{
    return (unsigned)c0 < (unsigned)b.dims[0] && (unsigned)c1 < (unsigned)b.dims[1] &&
           (unsigned)c2 < (unsigned)b.dims[2];
}

// Visit hash, identical to the oracle's: 64-bit mix over
// (brick id, i, j, k, bits(enter_t), bits(exit_t)).
__device__ __forceinline__ uint64_t vmix(uint64_t h, uint64_t w)
{
    h ^= w;
    h *= 0x100000001b3ULL;
    h ^= h >> 29;
    return h;
}

struct Rec {
    int64_t visits;
    int64_t samples;
    uint64_t hash;
    int brick;
    __device__ __forceinline__ void visit(double en, double ex, int c0, int c1, int c2)
    {
        uint64_t h = hash;
        h = vmix(h, (uint64_t)(uint32_t)brick | ((uint64_t)(uint32_t)c0 << 32));
        h = vmix(h, (uint64_t)(uint32_t)c1 | ((uint64_t)(uint32_t)c2 << 32));
        h = vmix(h, (uint64_t)__double_as_longlong(en));
        h = vmix(h, (uint64_t)__double_as_longlong(ex));
        hash = h;
        visits += 1;
    }
};

}  // namespace ytrc
