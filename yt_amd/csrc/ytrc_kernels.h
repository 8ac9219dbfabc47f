// ytrc_kernels.h -- the ray-casting kernel templates of libytrc.so.
//
// Included by the translation units that instantiate them (ytrc_k_*.hip, one
// per family of kernel variants, compiled in parallel) and by ytrc.hip for
// the shared constants.  Everything is in an unnamed namespace: each unit
// instantiates and exports (through ytrc_pick.h) only its own variants.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ytrc.h"
#include "ytrc_device.h"

namespace {

using namespace ytrc;


template <int NFP>
struct FieldLoad;
template <>
struct FieldLoad<1> {
    static __device__ __forceinline__ void get(const float *p, float v[1]) { v[0] = p[0]; }
};
template <>
struct FieldLoad<2> {
    static __device__ __forceinline__ void get(const float *p, float v[2])
    {
        float2 t = *reinterpret_cast<const float2 *>(p);
        v[0] = t.x;
        v[1] = t.y;
    }
};
template <>
struct FieldLoad<4> {
    static __device__ __forceinline__ void get(const float *p, float v[4])
    {
        float4 t = *reinterpret_cast<const float4 *>(p);
        v[0] = t.x;
        v[1] = t.y;
        v[2] = t.z;
        v[3] = t.w;
    }
};
template <>
struct FieldLoad<8> {
    static __device__ __forceinline__ void get(const float *p, float v[8])
    {
        float4 a = *reinterpret_cast<const float4 *>(p);
        float4 b = *reinterpret_cast<const float4 *>(p + 4);
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
        v[3] = a.w;
        v[4] = b.x;
        v[5] = b.y;
        v[6] = b.z;
        v[7] = b.w;
    }
};

__device__ __forceinline__ void kahan_add(float &s, float &c, float x)
{
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// ProjectionSampler.sample (image_samplers.pyx:330-343): rgba[i] +=
// data[i][cell] * (exit_t - enter_t) for i < min(4, n_fields).  No mask.
template <int NFP, bool REC>
struct ProjSampler {
    const float *data;
    int d1, d2;
    int dims0;
    int nbj, nbk;  // 4-cell blocks along axes 1 and 2
    float s[NFP];
    float c[NFP];
    Rec *rec;
};

// walk_volume's loop (grid_traversal.pyx:120-138) fused with
// ProjectionSampler.sample and software-pipelined: the DDA never depends
// on the data, so each round takes up to PF steps, issues their PF loads
// together and then accumulates them in visit order (the same additions, in
// the same order, as one visit at a time).
template <int NFP, bool REC>
__device__ __forceinline__ void walk_projection(const DBrick &b, Walk &w, double max_t,
                                                ProjSampler<NFP, REC> &S)
{
    constexpr int PF = 12;
    double t0 = w.tmax[0], t1 = w.tmax[1], t2 = w.tmax[2];
    int c0 = w.cur[0], c1 = w.cur[1], c2 = w.cur[2];
    double enter_t = w.enter_t;
    const int st0 = w.step[0], st1 = w.step[1], st2 = w.step[2];
    const double td0 = w.tdelta[0], td1 = w.tdelta[1], td2 = w.tdelta[2];
    bool live = true;
    while (live) {
        int off[PF];
        float dl[PF];
#pragma unroll
        for (int u = 0; u < PF; u++) {
            off[u] = -1;
            dl[u] = 0.0f;
            if (live) {
                const bool a01 = t0 < t1;
                const bool a02 = t0 < t2;
                const bool a12 = t1 < t2;
                const bool s0 = a01 & a02;
                const bool s1 = !a01 & a12;
                const bool s2 = !(s0 | s1);
                const double tm = s0 ? t0 : (s1 ? t1 : t2);
                const double exit_t = y_fmin(tm, max_t);
                if (REC) {
                    S.rec->visit(enter_t, exit_t, c0, c1, c2);
                    if (exit_t > enter_t) S.rec->samples += 1;
                }
                if ((unsigned)c0 < (unsigned)S.dims0 && (unsigned)c1 < (unsigned)S.d1 &&
                    (unsigned)c2 < (unsigned)S.d2) {
                    // 4x4x4-blocked cell offset (blocked_offset); block
                    // counts per face are below 2^24, so 24-bit multiplies
                    // are exact here
                    const unsigned blk =
                        __umul24(__umul24((unsigned)c0 >> 2, (unsigned)S.nbj) + ((unsigned)c1 >> 2),
                                 (unsigned)S.nbk) + ((unsigned)c2 >> 2);
                    off[u] = (int)((blk << 6) | ((c0 & 3) << 4) | ((c1 & 3) << 2) | (c2 & 3));
                    dl[u] = (float)(exit_t - enter_t);
                }
                // Step the chosen axis and test only its new index, as the
                // reference does, with lane masks rather than an axis-indexed
                // table (which the compiler would place in scratch memory).
                c0 += s0 ? st0 : 0;
                c1 += s1 ? st1 : 0;
                c2 += s2 ? st2 : 0;
                t0 = s0 ? t0 + td0 : t0;
                t1 = s1 ? t1 + td1 : t1;
                t2 = s2 ? t2 + td2 : t2;
                const bool out = (s0 & ((unsigned)c0 >= (unsigned)S.dims0)) |
                                 (s1 & ((unsigned)c1 >= (unsigned)S.d1)) |
                                 (s2 & ((unsigned)c2 >= (unsigned)S.d2));
                enter_t = tm;
                live = !(out || enter_t >= max_t);
            }
        }
        float v[PF][NFP];
#pragma unroll
        for (int u = 0; u < PF; u++) {
            if (off[u] >= 0) {
                FieldLoad<NFP>::get(S.data + (size_t)off[u] * NFP, v[u]);
            } else {
#pragma unroll
                for (int f = 0; f < NFP; f++) v[u][f] = 0.0f;
            }
        }
#pragma unroll
        for (int u = 0; u < PF; u++) {
            if (off[u] >= 0) {
#pragma unroll
                for (int f = 0; f < NFP; f++) kahan_add(S.s[f], S.c[f], v[u][f] * dl[u]);
            }
        }
    }
}

// Loads of one point's NFP fields through a raw buffer resource: an offset
// at or beyond the resource's byte count reads 0 (the hardware's range
// check), so a pipelined walk gives the steps past a ray's end the offset
// 0xffffffff instead of a branch around their loads.
template <int NFP>
struct BufLoad;
template <>
struct BufLoad<1> {
    static __device__ __forceinline__ void get(__amdgpu_buffer_rsrc_t r, unsigned off, float v[1])
    {
        v[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    }
};
template <>
struct BufLoad<2> {
    static __device__ __forceinline__ void get(__amdgpu_buffer_rsrc_t r, unsigned off, float v[2])
    {
        const auto t = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
        v[0] = __uint_as_float(t[0]);
        v[1] = __uint_as_float(t[1]);
    }
};
template <>
struct BufLoad<4> {
    static __device__ __forceinline__ void get(__amdgpu_buffer_rsrc_t r, unsigned off, float v[4])
    {
        const auto t = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
#pragma unroll
        for (int f = 0; f < 4; f++) v[f] = __uint_as_float(t[f]);
    }
};

// The smaller of two FP64 values that are not NaN, as one instruction
// (fmin's lowering first quiets each operand).
__device__ __forceinline__ double min_f64(double a, double b)
{
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// A raw buffer resource (gfx9 layout: base, stride 0, byte count, and the
// third dword's 32-bit data format) over `bytes` bytes at p.  Both are the
// same in every lane (one brick per wave) and are made scalar, so that the
// resource sits in SGPRs.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t brick_rsrc(const float *p, unsigned bytes)
{
    const uint64_t a = (uint64_t)p;
    const uint64_t lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)a);
    const uint64_t hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(a >> 32));
    return __builtin_amdgcn_make_buffer_rsrc((void *)((hi << 32) | lo), (short)0,
                                             __builtin_amdgcn_readfirstlane((int)bytes), 0x00020000);
}

// Per-wave LDS table of walk_projection_fast<NFP, true>: the offset terms of
// up to PTAB cells per axis, each axis's run with PTPAD padding words before
// and after.  A lane past its ray's end keeps stepping until its round ends
// (at most PTPAD - 1 more steps), so its indices stay within PTPAD of the
// brick and its reads within the wave's table.
constexpr int PTAB = 512;
constexpr int PTPAD = 12;
constexpr int PTAB0 = PTPAD, PTAB1 = PTAB + 3 * PTPAD, PTAB2 = 2 * PTAB + 5 * PTPAD;
constexpr int PTAB_WAVE = 3 * PTAB + 6 * PTPAD;  // words per wave

// walk_projection for the common case, taken per wave and brick when
// every walking lane starts with all three cell indices inside the brick
// (so that, stepping one axis at a time and stopping when that axis leaves
// the brick, every visit is inside it), the brick's points span fewer than
// 2^32 bytes and its 4x4x4-block strides fit 24-bit multiplies:
// * no per-visit bounds test (walk_projection keeps it for rays that start
//   within 1e-10 dds outside a face, whose indices stay out of range on the
//   axis they never step);
// * the blocked offset as 16 c0 + 4 c1 + c2 + (c0>>2) K0 + (c1>>2) K1 +
//   (c2>>2) 60, with K0 = 64 (nbj nbk - 1), K1 = 64 nbk - 16 (the same
//   value as blocked_offset, in fewer operations);
// * the steps of a round run for every lane of the wave; a lane past its
//   ray's end gives its load the out-of-range offset, which reads 0 (its
//   lengths stay finite, so its products are 0);
// * each round's visits are summed in visit order with fused multiply-adds
//   and the round's sum is added to the image with Kahan compensation.
// TAB (persistent launches, brick sides up to PTAB): the offset is a sum
// of three per-axis terms, 4 NFP ((c0>>2) K0 + 16 c0), 4 NFP ((c1>>2) K1 +
// 4 c1) and 4 NFP ((c2>>2) 60 + c2), which the wave tabulates in LDS (tab,
// PTAB_WAVE words) for the brick's sides before it walks; a step then
// reads the new term of the axis it steps instead of forming the offset.
template <int NFP, bool TAB = false>
__device__ __forceinline__ void walk_projection_fast(Walk &w, double max_t, ProjSampler<NFP, false> &S,
                                                     __amdgpu_buffer_rsrc_t rsrc, unsigned K0, unsigned K1,
                                                     unsigned *tab = nullptr)
{
    constexpr int PF = 12;
    static_assert(PF <= PTPAD, "a round's steps past a ray's end must stay in the table's padding");
    double t0 = w.tmax[0], t1 = w.tmax[1], t2 = w.tmax[2];
    int c0 = w.cur[0], c1 = w.cur[1], c2 = w.cur[2];
    double enter_t = w.enter_t;
    const int st0 = w.step[0], st1 = w.step[1], st2 = w.step[2];
    const double td0 = w.tdelta[0], td1 = w.tdelta[1], td2 = w.tdelta[2];
    bool live = true;
    unsigned f0 = 0, f1 = 0, f2 = 0;  // (TAB) the three terms at (c0, c1, c2)
    if constexpr (TAB) {
        // the active lanes fill the terms of the brick's cells in turn
        const uint64_t act = __ballot(1);
        const int nact = __popcll(act);
        const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(act >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((unsigned)act, 0u));
        constexpr unsigned sc = 4u * NFP;
        for (int i = rank; i < S.dims0; i += nact)
            tab[PTAB0 + i] = (__umul24((unsigned)i >> 2, K0) + ((unsigned)i << 4)) * sc;
        for (int i = rank; i < S.d1; i += nact)
            tab[PTAB1 + i] = (__umul24((unsigned)i >> 2, K1) + ((unsigned)i << 2)) * sc;
        for (int i = rank; i < S.d2; i += nact)
            tab[PTAB2 + i] = (__umul24((unsigned)i >> 2, 60u) + (unsigned)i) * sc;
        __builtin_amdgcn_wave_barrier();
        f0 = tab[PTAB0 + c0];
        f1 = tab[PTAB1 + c1];
        f2 = tab[PTAB2 + c2];
    }
    while (live) {
        unsigned off[PF];
        float dl[PF];
#pragma unroll
        for (int u = 0; u < PF; u++) {
            // The reference's axis choice (tmax[0] < tmax[1] ? (tmax[0] <
            // tmax[2] ? 0 : 2) : (tmax[1] < tmax[2] ? 1 : 2)) from two
            // comparisons: m = the first pair's choice, then m against t2.
            // The chosen tmax is the smallest (the tmax are never NaN), so
            // m and tm are minima, one instruction each.
            const bool a01 = t0 < t1;
            const double m = min_f64(t0, t1);
            const bool bm = m < t2;
            const bool s0 = a01 & bm;
            const bool s1 = !a01 & bm;
            const double tm = min_f64(m, t2);
            // exit_t = fmin(tm, max_t), and the loop's enter_t >= max_t
            // test on the next enter_t (= tm), from one comparison
            const bool ge = tm >= max_t;
            const double exit_t = min_f64(tm, max_t);
            if constexpr (TAB) {
                off[u] = live ? f0 + f1 + f2 : 0xffffffffu;
                // formed at its step, like the length: no step's terms stay
                // live until the round's loads
                asm volatile("" : "+v"(off[u]));
            } else {
                const unsigned lin = ((unsigned)c0 << 4) + ((unsigned)c1 << 2) + (unsigned)c2;
                const unsigned o = __umul24((unsigned)c0 >> 2, K0) +
                                   (__umul24((unsigned)c1 >> 2, K1) + (__umul24((unsigned)c2 >> 2, 60u) + lin));
                off[u] = live ? o * (unsigned)(4 * NFP) : 0xffffffffu;
            }
            // (a lane past its end reads 0; its length stays finite)
            dl[u] = (float)(exit_t - enter_t);
            // the length is formed here, so that no step's FP64 times stay
            // live until the round's loads
            asm volatile("" : "+v"(dl[u]));
            // Step the chosen axis under an execution mask (the empty asm
            // keeps each update a masked region rather than selects).
            // (a term read at a step out of the brick lands in the table's
            // padding and is never used: the lane is past its end)
            if (s0) {
                asm volatile("");
                c0 += st0;
                t0 += td0;
                if constexpr (TAB) f0 = tab[PTAB0 + c0];
            }
            if (s1) {
                asm volatile("");
                c1 += st1;
                t1 += td1;
                if constexpr (TAB) f1 = tab[PTAB1 + c1];
            }
            const bool s2 = !(s0 | s1);
            if (s2) {
                asm volatile("");
                c2 += st2;
                t2 += td2;
                if constexpr (TAB) f2 = tab[PTAB2 + c2];
            }
            const bool out = (s0 & ((unsigned)c0 >= (unsigned)S.dims0)) |
                             (s1 & ((unsigned)c1 >= (unsigned)S.d1)) |
                             (s2 & ((unsigned)c2 >= (unsigned)S.d2));
            enter_t = tm;
            live = live & !(out | ge);
        }
        float v[PF][NFP];
#pragma unroll
        for (int u = 0; u < PF; u++) BufLoad<NFP>::get(rsrc, off[u], v[u]);
        float r[NFP];
#pragma unroll
        for (int f = 0; f < NFP; f++) r[f] = 0.0f;
#pragma unroll
        for (int u = 0; u < PF; u++)
#pragma unroll
            for (int f = 0; f < NFP; f++) r[f] = __builtin_fmaf(v[u][f], dl[u], r[f]);
#pragma unroll
        for (int f = 0; f < NFP; f++) kahan_add(S.s[f], S.c[f], r[f]);
    }
}

template <int N>
__device__ __forceinline__ float pick(const float (&a)[N], int i)
{
    float r = a[0];
#pragma unroll
    for (int k = 1; k < N; k++) r = (i == k) ? a[k] : r;
    return r;
}

// Transfer-function parameters of one image, loaded once per tile into
// (uniform) registers.  mode 1 is the ColorTransferFunction pattern
// (transfer_functions.py:385-408): four tables on one field with shared
// bounds, red/green/blue weighted by the alpha table, channels linked 0..3;
// in transmittance form its LDS table holds per bin {y - k dy} and {dy} of
// the four tables (fit_eval_ctf).
// mode 0 is the general FieldInterpolationTable set.
struct TFCache {
    int mode;
    int n_fits;
    int grey;
    int fid;       // mode 1: the one field
    float b0, b1, idbin, nb0idbin;  // nb0idbin = -b0 * idbin
    float qc1, qzb;                 // transmittance-form CTF entry selection
    int nbm2;      // mode 1: nbins - 2
    int nbins[6];
    int off[6];
    int field_id[6];
    int wfield[6];
    int chexp[4];
    float tb0[6], tb1[6], tidbin[6];
};

__device__ __forceinline__ float unif(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Loads the (wave-uniform) table parameters into scalar registers.
__device__ __forceinline__ void load_tf(const DTF &T, TFCache &c)
{
    c.mode = uni(T.mode);
    c.n_fits = uni(T.n_fits);
    c.grey = uni(T.grey);
    c.fid = uni(T.field_id[0]);
    c.b0 = unif(T.b0[0]);
    c.b1 = unif(T.b1[0]);
    c.idbin = unif(T.idbin[0]);
    c.nb0idbin = -c.b0 * c.idbin;
    c.qc1 = unif(T.qc1);
    c.qzb = unif(T.qzb);
    c.nbm2 = uni(T.nbins[0] - 2);
#pragma unroll
    for (int t = 0; t < 6; t++) {
        c.nbins[t] = uni(T.nbins[t]);
        c.off[t] = uni(T.off[t]);
        c.field_id[t] = uni(T.field_id[t]);
        c.wfield[t] = uni(T.wfield[t]);
        c.tb0[t] = unif(T.b0[t]);
        c.tb1[t] = unif(T.b1[t]);
        c.tidbin[t] = unif(T.idbin[t]);
    }
#pragma unroll
    for (int ch = 0; ch < 4; ch++) c.chexp[ch] = uni(T.chexp[ch]);
}

// FIT_get_value + FIT_eval_transfer (field_interpolation_tables.pxd:66-110),
// general table set, tables in LDS as float4 {y, dy, d0, 0}.  The
// reference's sequential weight-table products (istorage[i] *=
// istorage[weight_table_id]) and channel links (trgba[c] =
// istorage[field_table_ids[c]]) are resolved on the host into chexp[c]:
// 2 bits per raw table u giving the power of table u's value in channel c.
// Table reads are branch-free: out-of-range values read a clamped bin and
// are then zeroed.
template <int NFP>
__device__ __forceinline__ void fit_eval_general(const TFCache &tf, const float4 *__restrict__ lt,
                                                 const float (&dvs)[NFP], float e[4])
{
    float raw[6];
#pragma unroll
    for (int t = 0; t < 6; t++) {
        raw[t] = 0.0f;
        if (t < tf.n_fits) {
            const float v = (NFP == 1) ? dvs[0] : pick(dvs, tf.field_id[t]);
            const bool ok = (v < tf.tb1[t]) && (v > tf.tb0[t]) && __builtin_isnormal(v);
            int bin = (int)((v - tf.tb0[t]) * tf.tidbin[t]);
            bin = bin < 0 ? 0 : (bin > tf.nbins[t] - 2 ? tf.nbins[t] - 2 : bin);
            const float4 q = lt[tf.off[t] + bin];
            float r = __builtin_fmaf(v - q.z, q.y, q.x);
            if (tf.wfield[t] != -1) r *= (NFP == 1) ? dvs[0] : pick(dvs, tf.wfield[t]);
            raw[t] = ok ? r : 0.0f;
        }
    }
#pragma unroll
    for (int ch = 0; ch < 4; ch++) {
        const int ce = tf.chexp[ch];
        float r = (ce < 0) ? 0.0f : 1.0f;
#pragma unroll
        for (int u = 0; u < 6; u++) {
            const int ex = (ce >> (2 * u)) & 3;
            if (ex >= 1) r *= raw[u];
            if (ex >= 2) r *= raw[u];
            if (ex >= 3) r *= raw[u];
        }
        e[ch] = r;
    }
}

// The ColorTransferFunction case of the same evaluation: one bin for the
// four tables, e_c = table_c(v) * table_alpha(v) for c < 3, e_3 = alpha.
// An out-of-range or non-normal v reads bin 0 and gets alpha 0, so every
// channel is 0 without a select per channel.
// ColorTransferFunction emissions in the transmittance-form layout: entry
// k = floor(u) (48 B apart) holds {y - k dy} and {dy} of the red, green,
// blue and alpha tables, so that each lerp y + (u - k) dy is one FMA in the bin coordinate
// u = 1 + (v - b0) * idbin, and e_c = lerp_c * lerp_alpha.  Two 16-byte LDS
// reads per sub-sample: the loop is bound by LDS issue as much as by VALU.
// FIT_get_value's range test v in (b0, b1)
// (field_interpolation_tables.pxd:63-82) becomes u in [1, nb) through the
// zero entries 0 and nb, nb + 1 and a clamp of u to [0, nb + 1]; the test
// is thereby made on the FP32 bin coordinate, exact up to its rounding.
// FIT_get_value's test for values that are not normal is made once per
// visit by the caller (quad_visit_ok), so v is finite.  The caller passes
// u unclamped: along a visit's sub-samples it is the cubic in s whose
// coefficients are the value's times idbin, plus qc1 on the constant one.
__device__ __forceinline__ void fit_eval_ctf(const TFCache &tf, const float4 *__restrict__ lt,
                                             float u, float e[3])
{
    u = __builtin_amdgcn_fmed3f(u, 0.0f, tf.qzb);
    const unsigned bin = (unsigned)u;
    const float4 *q = (const float4 *)((const char *)lt + __umul24(bin, 48u));
    const float4 y = q[0];
    const float4 d = q[1];
    const float a = __builtin_fmaf(d.w, u, y.w);
    e[0] = __builtin_fmaf(d.x, u, y.x) * a;
    e[1] = __builtin_fmaf(d.y, u, y.y) * a;
    e[2] = __builtin_fmaf(d.z, u, y.z) * a;
}

// Brick fields the samplers read, copied by value (uniform across a wave).
struct BrickReg {
    double le[3];
    double dds[3];
    double idds[3];
    int d0, d1, d2;
    int vnbj, vnbk;  // vertex blocks along axes 1 and 2
};

__device__ __forceinline__ BrickReg brick_reg(const DBrick &b)
{
    BrickReg r;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        r.le[i] = uni_d(b.le[i]);
        r.dds[i] = uni_d(b.dds[i]);
        r.idds[i] = uni_d(b.idds[i]);
    }
    r.d0 = uni(b.dims[0]);
    r.d1 = uni(b.dims[1]);
    r.d2 = uni(b.dims[2]);
    r.vnbj = vblocks(r.d1);
    r.vnbk = vblocks(r.d2);
    return r;
}

// Offset (in points) of corner (0, 0, 0) of cell (c0, c1, c2) of a vertex
// brick; the 24-bit multiplies are exact because a brick has fewer than
// 2^24 vertex blocks per face (checked at upload).
__device__ __forceinline__ int vertex_base(const BrickReg &b, int c0, int c1, int c2)
{
    const unsigned blk = __umul24(__umul24((unsigned)c0 >> VB_SHIFT, (unsigned)b.vnbj) +
                                      ((unsigned)c1 >> VB_SHIFT),
                                  (unsigned)b.vnbk) +
                         ((unsigned)c2 >> VB_SHIFT);
    return (int)((blk << VBP_SHIFT) +
                 __umul24(__umul24((unsigned)c0 & (VB - 1), VB1) + ((unsigned)c1 & (VB - 1)), VB1) +
                 ((unsigned)c2 & (VB - 1)));
}

// The 8 corner values of a cell, fields interleaved, from its corner
// (0, 0, 0) at p0: cv[4a + 2b + c] is corner (a, b, c).
template <int NFP>
__device__ __forceinline__ void load_corners(const float *p0, float (&cv)[8][NFP])
{
    FieldLoad<NFP>::get(p0, cv[0]);
    FieldLoad<NFP>::get(p0 + vcorner(0, 0, 1) * NFP, cv[1]);
    FieldLoad<NFP>::get(p0 + vcorner(0, 1, 0) * NFP, cv[2]);
    FieldLoad<NFP>::get(p0 + vcorner(0, 1, 1) * NFP, cv[3]);
    FieldLoad<NFP>::get(p0 + vcorner(1, 0, 0) * NFP, cv[4]);
    FieldLoad<NFP>::get(p0 + vcorner(1, 0, 1) * NFP, cv[5]);
    FieldLoad<NFP>::get(p0 + vcorner(1, 1, 0) * NFP, cv[6]);
    FieldLoad<NFP>::get(p0 + vcorner(1, 1, 1) * NFP, cv[7]);
}

// offset_interpolate (fixed_interpolator.cpp:34-52) restricted to the
// sub-sample positions u(s) = dp + s*ds of one visit is a cubic in s.  Its
// coefficients come from the 8 corners c[4i+2j+k] once per visit; each
// sub-sample then costs 3 FMAs.
__device__ __forceinline__ void trilinear_cubic(const float (&c)[8], const float dp[3],
                                                const float ds[3], float k[4])
{
    const float A = c[0];
    const float Bz = c[1] - c[0];
    const float By = c[2] - c[0];
    const float Bx = c[4] - c[0];
    const float Cyz = (c[3] - c[2]) - Bz;
    const float Cxz = (c[5] - c[4]) - Bz;
    const float Cxy = (c[6] - c[4]) - By;
    const float D = ((c[7] - c[6]) - (c[5] - c[4])) - Cyz;
    const float x0 = dp[0], y0 = dp[1], z0 = dp[2];
    const float dx = ds[0], dy = ds[1], dz = ds[2];
    const float fxy = __builtin_fmaf(D, z0, Cxy);
    const float fxz = __builtin_fmaf(D, y0, Cxz);
    const float fyz = __builtin_fmaf(D, x0, Cyz);
    const float tyz = __builtin_fmaf(Cyz, z0, By);
    const float ex = __builtin_fmaf(y0, fxy, __builtin_fmaf(Cxz, z0, Bx));
    const float ey = __builtin_fmaf(x0, fxy, tyz);
    const float ez = __builtin_fmaf(x0, fxz, __builtin_fmaf(Cyz, y0, Bz));
    k[0] = __builtin_fmaf(x0, ex, __builtin_fmaf(y0, tyz, __builtin_fmaf(z0, Bz, A)));
    k[1] = __builtin_fmaf(ex, dx, __builtin_fmaf(ey, dy, ez * dz));
    const float dydz = dy * dz;
    k[2] = __builtin_fmaf(fxy, dx * dy, __builtin_fmaf(fxz, dx * dz, fyz * dydz));
    k[3] = D * dx * dydz;
}

// FIT_get_value returns 0 for values that are not normal
// (field_interpolation_tables.pxd:74, isnormal: false for 0, subnormals, inf
// and NaN).  The ColorTransferFunction path makes that test once per visit
// on the cubic k0 + k1 s + k2 s^2 + k3 s^3 of the sampled field:
// * a NaN or infinite corner makes every sub-sample value of the visit
//   non-finite (k0 depends on all 8 corners, and 0 * inf is NaN), and makes
//   |k0| + |k1| + |k2| + |k3| non-finite;
// * all coefficients 0 is a value of exactly 0 at every sub-sample.
// Either way every sub-sample gives 0 and the visit changes nothing, so it
// is skipped.  Otherwise the values are finite, and a sub-sample value that
// is 0 or subnormal only through FP32 cancellation is looked up like its
// neighbours (the FP64 value it approximates is normal).  Coefficients whose
// magnitudes sum beyond FLT_MAX count as non-finite.
__device__ __forceinline__ bool quad_visit_ok(const float (&k)[4])
{
    const float m = __builtin_fabsf(k[0]) + __builtin_fabsf(k[1]) + __builtin_fabsf(k[2]) +
                    __builtin_fabsf(k[3]);
    return __builtin_amdgcn_classf(m, (1 << 7) | (1 << 8));  // +subnormal or +normal
}

// The sub-samples of one visit for a ColorTransferFunction in transmittance
// form.  ku: the TF bin coordinate along the visit's sub-samples, u(s) =
// ku0 + ku1 s + ku2 s^2 + ku3 s^3; dtf: the sub-sample length.  The visit's
// opacity Q_c = 1 - prod_s (1 - dt e_s) by the recurrence Q' = Q + e G,
// G = dt (1 - Q): Q stays small where the medium is thin, so FP32 keeps its
// relative precision; the FP64 transmittance st then takes T (1 - Q) once
// per visit.
template <int NS>
__device__ __forceinline__ void ctf_visit(const TFCache &tf, const float4 *__restrict__ lt,
                                          const float (&ku)[4], float dtf, int ns, double (&st)[4])
{
    float Q[3] = {0.0f, 0.0f, 0.0f};
    float G[3] = {dtf, dtf, dtf};
    const int nsq = (NS > 0) ? NS : ns;
    float sf = 0.0f;
#pragma unroll
    for (int s = 0; s < nsq; s++, sf += 1.0f) {
        float e[3];
        // Sub-sample 0 is ku0 (the coefficients are finite here).
        const float u = (NS > 0 && s == 0)
                            ? ku[0]
                            : __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(ku[3], sf, ku[2]), sf, ku[1]),
                                             sf, ku[0]);
        fit_eval_ctf(tf, lt, u, e);
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            Q[ch] = __builtin_fmaf(e[ch], G[ch], Q[ch]);
            G[ch] = __builtin_fmaf(-dtf, Q[ch], dtf);
        }
    }
#pragma unroll
    for (int ch = 0; ch < 3; ch++) st[ch] = __builtin_fma(-st[ch], (double)Q[ch], st[ch]);
}

// VolumeRenderSampler.sample (image_samplers.pyx:453-487).
// MODE bit 0: ColorTransferFunction in transmittance form; bit 1: n_samples == 10 unrolled;
// bit 2: early ray termination.
// st[] is the compositing state: rgba, or, when `tform` is set (non-grey
// and the host has shown dt * e <= 1 for every sub-sample), the per-channel
// transmittance T = 1 - rgba, for which rgba' = a + max(1 - a, 0) rgba
// (field_interpolation_tables.pxd:107-110) is exactly T' = T (1 - a).
template <int NFP, bool REC, int MODE>
struct VRSampler {
    const float *data;
    const uint8_t *mask;
    BrickReg b;
    double pa[3];  // v_dir * idds (wave-uniform)
    float paf[3];  // (float)pa
    double pb[3];  // (v_pos - left_edge) * idds
    const TFCache *tf;
    const float4 *lt;
    int nb;
    int tform;
    static constexpr bool ert = (MODE & 4) != 0;  // also track the brick's transmittance tb
    double inv_ns;
    int n_samples;
    double st[4];
    double tb[4];
    Rec *rec;

    __device__ __forceinline__ void eval(const float (&dvs)[NFP], float e[4]) const
    {
        fit_eval_general<NFP>(*tf, lt, dvs, e);
    }

    template <int NS, class Body>
    __device__ __forceinline__ static void each_sample(int ns, Body &&body)
    {
        if constexpr (NS > 0) {
#pragma unroll
            for (int s = 0; s < NS; s++) body(s);
        } else {
            for (int s = 0; s < ns; s++) body(s);
        }
    }

    template <int NS>
    __device__ __forceinline__ void ctf_samples(const float (&ku)[4], double dt, int ns)
    {
        ctf_visit<NS>(*tf, lt, ku, (float)dt, ns, st);
    }

    template <int NS>
    __device__ __forceinline__ void samples(const float (&k)[NFP][4], double dt, int ns)
    {
        const TFCache &T = *tf;
        auto value = [&](int s, float (&dvs)[NFP]) {
            const float sf = (float)s;
#pragma unroll
            for (int f = 0; f < NFP; f++)
                dvs[f] = __builtin_fmaf(
                    __builtin_fmaf(__builtin_fmaf(k[f][3], sf, k[f][2]), sf, k[f][1]), sf, k[f][0]);
        };
        if (tform) {
            each_sample<NS>(ns, [&](int s) {
                float dvs[NFP], e[4];
                value(s, dvs);
                eval(dvs, e);
#pragma unroll
                for (int ch = 0; ch < 3; ch++) st[ch] *= __builtin_fma(-dt, (double)e[ch], 1.0);
            });
        } else if (T.grey) {
            each_sample<NS>(ns, [&](int s) {
                float dvs[NFP], e[4];
                value(s, dvs);
                eval(dvs, e);
                const double ta = y_fmax(1.0 - dt * (double)e[3], 0.0);
#pragma unroll
                for (int ch = 0; ch < 4; ch++) st[ch] = __builtin_fma(ta, st[ch], dt * (double)e[ch]);
                if constexpr (ert) tb[0] *= ta;
            });
        } else {
            each_sample<NS>(ns, [&](int s) {
                float dvs[NFP], e[4];
                value(s, dvs);
                eval(dvs, e);
#pragma unroll
                for (int ch = 0; ch < 3; ch++) {
                    const double a = dt * (double)e[ch];
                    const double ta = y_fmax(1.0 - a, 0.0);
                    st[ch] = __builtin_fma(ta, st[ch], a);
                    if constexpr (ert) tb[ch] *= ta;
                }
            });
        }
    }

    __device__ __forceinline__ void operator()(double en, double ex, int c0, int c1, int c2)
    {
        if (REC) rec->visit(en, ex, c0, c1, c2);
        const int d1 = b.d1, d2 = b.d2;
        if (!((unsigned)c0 < (unsigned)b.d0 && (unsigned)c1 < (unsigned)d1 &&
              (unsigned)c2 < (unsigned)d2))
            return;
        const int cell = (int)__umul24(__umul24(c0, d1) + c1, d2) + c2;
        const int voff = vertex_base(b, c0, c1, c2);
        // The mask byte and the corner values are independent loads: issue
        // both first, then form the sample positions while they are in
        // flight (the barrier keeps the scheduler from moving that work
        // ahead of the loads; the empty asm below keeps the first use of a
        // corner after it).
        const int mval = (mask != nullptr) ? (int)mask[cell] : 1;
        float cv[8][NFP];
        load_corners<NFP>(data + (size_t)voff * NFP, cv);
        __builtin_amdgcn_sched_barrier(0);
        // dt = (exit_t - enter_t) / n_samples, as a product with 1/n_samples.
        const double dt = (ex - en) * inv_ns;
        const double tt = en + 0.5 * dt;
        // The reference's cell-relative position is
        // (tt*v_dir + v_pos - (idx*dds + LE)) * idds (image_samplers.pyx:
        // 475-479).  Here it is refactored as fma(tt, pa, pb) - idx with
        // pa = v_dir*idds and pb = (v_pos - LE)*idds formed once per brick:
        // the same value up to a few FP64 roundings, far below the FP32 it
        // is converted to.
        // The sub-sample step v_dir * idds * dt in FP32 (paf = (float)pa,
        // per brick): it only spaces the sub-samples within the cell.
        const float dtf = (float)dt;
        float dp[3], ds[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int ci = (i == 0) ? c0 : ((i == 1) ? c1 : c2);
            dp[i] = (float)(__builtin_fma(tt, pa[i], pb[i]) - (double)ci);
            ds[i] = paf[i] * dtf;
        }
        // The positions are formed before the corner values are first used.
        asm volatile("" : "+v"(dp[0]), "+v"(dp[1]), "+v"(dp[2]), "+v"(ds[0]), "+v"(ds[1]), "+v"(ds[2]));
#pragma unroll
        for (int q = 0; q < 8; q++)
#pragma unroll
            for (int f = 0; f < NFP; f++) asm volatile("" : "+v"(cv[q][f]));
        if (mval != 1) return;
        if (REC) rec->samples += n_samples;
        float k[NFP][4];
#pragma unroll
        for (int f = 0; f < NFP; f++) {
            const float cf[8] = {cv[0][f], cv[1][f], cv[2][f], cv[3][f],
                                 cv[4][f], cv[5][f], cv[6][f], cv[7][f]};
            trilinear_cubic(cf, dp, ds, k[f]);
        }
        if constexpr ((MODE & 1) != 0) {
            float kq[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                kq[j] = k[0][j];
#pragma unroll
                for (int f = 1; f < NFP; f++) kq[j] = (tf->fid == f) ? k[f][j] : kq[j];
            }
            if (!quad_visit_ok(kq)) return;
            // The bin coordinate u = v idbin + qc1 of the value's cubic.
            const float ku[4] = {__builtin_fmaf(kq[0], tf->idbin, tf->qc1), kq[1] * tf->idbin,
                                 kq[2] * tf->idbin, kq[3] * tf->idbin};
            if (MODE & 2)
                ctf_samples<10>(ku, dt, 10);
            else
                ctf_samples<0>(ku, dt, n_samples);
            return;
        }
        if (MODE & 2)
            samples<10>(k, dt, 10);
        else
            samples<0>(k, dt, n_samples);
    }
};

// Byte offsets of the vertex blocks (vertex_offset times 4, one field) as
// a sum over axes: (i & 3) W = i W - (i >> 2) 4 W, so the offset of
// corner (0, 0, 0) of cell (c0, c1, c2) is
//   4 (25 c0 + 5 c1 + c2) + (c0 >> 2) K0 + (c1 >> 2) K1 + (c2 >> 2) K2
// with K0 = 4 (128 nbj nbk - 100), K1 = 4 (128 nbk - 20), K2 = 4 (128 - 4)
// for VB = 4, VB1 = 5, VBP = 128 (nbj, nbk: vertex blocks along axes 1, 2).
constexpr unsigned VW0 = 4u * VB1 * VB1, VW1 = 4u * VB1, VW2 = 4u;
constexpr unsigned VK2 = 4u * (VBP - VB);
__host__ __device__ constexpr uint64_t vk0_bytes(uint64_t nbj, uint64_t nbk) { return 4u * (VBP * nbj * nbk - VB * VB1 * VB1); }
__host__ __device__ constexpr uint64_t vk1_bytes(uint64_t nbk) { return 4u * (VBP * nbk - VB * VB1); }

// VRSampler with a ColorTransferFunction in transmittance form (MODE bit
// 0), one field and a shared ray direction, walked by walk_loop's DDA, for
// the common case, taken per wave and brick when every walking lane starts
// with all three cell indices inside the brick (so every visit is inside
// it: one axis steps at a time and the walk stops when it leaves), the
// brick's vertex blocks span fewer than 2^32 bytes and K0, K1 fit 24-bit
// multiplies.  The same visits and sub-samples in fewer operations:
// * no per-visit bounds test;
// * corner (0, 0, 0)'s byte offset from the sum over axes above, the 8
//   corners read through a buffer resource (four 8-byte loads);
// * the cell-relative position fma(tt, pa, r) with r = pb - c kept per axis
//   and stepped with its index, tt = en + (ex - en) / (2 n_samples);
// * the cubic in the bin coordinate from the field's partial derivatives
//   and per-brick products of P = (float)(v_dir idds): with the
//   sub-sample length d, u(s) = idbin (k0 + (s d) (fx P0 + fy P1 + fz P2) +
//   (s d)^2 (fxy P0 P1 + fxz P0 P2 + fyz P1 P2) + (s d)^3 D P0 P1 P2) + qc1;
// * the axis-0/1 choice and the axis-2 step as nested branches, so that
//   the execution masks come from the two comparisons.
// VLDS (MODE bit 4, YTRC_BATCH_VERTEX_LDS) stages vertex blocks in LDS:
// each wave owns a 512-byte buffer (vbuf) holding one block of the brick.
// At a visit where every active lane is in the same block, the wave reads
// that block into vbuf once (the active lanes share its 64 8-byte pieces)
// and takes the 8 corners from LDS while it stays there; at other visits
// the lanes read their corners through the buffer resource.  The same
// values either way.
template <int NS, bool VLDS = false>
__device__ __forceinline__ void walk_vr_ctf_fast(Walk &w, double max_t, const BrickReg &b,
                                                 __amdgpu_buffer_rsrc_t rsrc, unsigned K0, unsigned K1,
                                                 const uint8_t *mask, const double (&pa)[3],
                                                 const double (&pb)[3], const TFCache &tf,
                                                 const float4 *__restrict__ lt, double inv_ns, int ns,
                                                 double (&st)[4], float *vbuf = nullptr)
{
    unsigned staged_at = 0xffffffffu;  // byte offset of the block in vbuf (wave-uniform)
    double t0 = w.tmax[0], t1 = w.tmax[1], t2 = w.tmax[2];
    int c0 = w.cur[0], c1 = w.cur[1], c2 = w.cur[2];
    double en = w.enter_t;
    // (the direction is shared: steps, increments and strides are scalar)
    const int st0 = uni(w.step[0]), st1 = uni(w.step[1]), st2 = uni(w.step[2]);
    const double td0 = uni_d(w.tdelta[0]), td1 = uni_d(w.tdelta[1]), td2 = uni_d(w.tdelta[2]);
    // the steps as doubles from scalar selects (no conversion at each step)
    const double sd0 = (st0 > 0) ? 1.0 : ((st0 < 0) ? -1.0 : 0.0);
    const double sd1 = (st1 > 0) ? 1.0 : ((st1 < 0) ? -1.0 : 0.0);
    const double sd2 = (st2 > 0) ? 1.0 : ((st2 < 0) ? -1.0 : 0.0);
    K0 = (unsigned)uni((int)K0);
    K1 = (unsigned)uni((int)K1);
    double r0 = pb[0] - (double)c0, r1 = pb[1] - (double)c1, r2 = pb[2] - (double)c2;
    const double hns = 0.5 * inv_ns;
    const float inv_nsf = unif((float)inv_ns);
    const float P0 = unif((float)pa[0]), P1 = unif((float)pa[1]), P2 = unif((float)pa[2]);
    const float P01 = unif(P0 * P1), P02 = unif(P0 * P2), P12 = unif(P1 * P2);
    const float P012 = unif(P01 * P2);
    const float idb = tf.idbin, qc1 = tf.qc1;
    const unsigned D0 = (unsigned)b.d0, D1 = (unsigned)b.d1, D2 = (unsigned)b.d2;
    while (true) {
        // walk_loop's axis choice and exit time (grid_traversal.pyx:120-138);
        // the chosen tmax is the smallest, so its value is a minimum (the
        // tmax are never NaN)
        const bool a01 = t0 < t1;
        const double m = min_f64(t0, t1);
        const bool bm = m < t2;
        const double tm = min_f64(m, t2);
        const bool ge = tm >= max_t;
        const double ex = min_f64(tm, max_t);
        {
            const unsigned o = __umul24((unsigned)c0 >> 2, K0) +
                               (__umul24((unsigned)c1 >> 2, K1) +
                                (__umul24((unsigned)c2 >> 2, VK2) +
                                 (__umul24((unsigned)c0, VW0) + (__umul24((unsigned)c1, VW1) + ((unsigned)c2 << 2)))));
            int mval = 1;
            if (mask != nullptr) mval = mask[(int)__umul24(__umul24(c0, D1) + c1, D2) + c2];
            float cv[8];
            bool from_lds = false;
            if constexpr (VLDS) {
                // the block's byte offset and the corner's offset inside it
                const unsigned ib = __umul24((unsigned)c0 & 3u, VW0) +
                                    (__umul24((unsigned)c1 & 3u, VW1) + (((unsigned)c2 & 3u) << 2));
                const unsigned bs = o - ib;
                const unsigned bs0 = (unsigned)__builtin_amdgcn_readfirstlane((int)bs);
                if (__all(bs == bs0)) {
                    if (bs0 != staged_at) {
                        const uint64_t act = __ballot(1);
                        const int nact = __popcll(act);
                        const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(act >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)act, 0u));
                        for (int k = rank; k < 64; k += nact) {
                            float v2[2];
                            BufLoad<2>::get(rsrc, bs0 + 8u * (unsigned)k, v2);
                            vbuf[2 * k] = v2[0];
                            vbuf[2 * k + 1] = v2[1];
                        }
                        __builtin_amdgcn_wave_barrier();
                        staged_at = bs0;
                    }
                    const float *q = vbuf + (ib >> 2);
                    cv[0] = q[vcorner(0, 0, 0)];
                    cv[1] = q[vcorner(0, 0, 1)];
                    cv[2] = q[vcorner(0, 1, 0)];
                    cv[3] = q[vcorner(0, 1, 1)];
                    cv[4] = q[vcorner(1, 0, 0)];
                    cv[5] = q[vcorner(1, 0, 1)];
                    cv[6] = q[vcorner(1, 1, 0)];
                    cv[7] = q[vcorner(1, 1, 1)];
                    from_lds = true;
                }
            }
            if (!from_lds) {
                BufLoad<2>::get(rsrc, o, cv);
                BufLoad<2>::get(rsrc, o + 4u * vcorner(0, 1, 0), cv + 2);
                BufLoad<2>::get(rsrc, o + 4u * vcorner(1, 0, 0), cv + 4);
                BufLoad<2>::get(rsrc, o + 4u * vcorner(1, 1, 0), cv + 6);
            }
            __builtin_amdgcn_sched_barrier(0);
            const double d = ex - en;
            const double tt = __builtin_fma(d, hns, en);
            float x0 = (float)__builtin_fma(tt, pa[0], r0);
            float y0 = (float)__builtin_fma(tt, pa[1], r1);
            float z0 = (float)__builtin_fma(tt, pa[2], r2);
            const float dtf = (float)d * inv_nsf;
            asm volatile("" : "+v"(x0), "+v"(y0), "+v"(z0));
#pragma unroll
            for (int q = 0; q < 8; q++) asm volatile("" : "+v"(cv[q]));
            if (mval == 1) {
                // offset_interpolate's trilinear form (fixed_interpolator.cpp:
                // 34-52) as A + Bx x + By y + Bz z + Cxy xy + Cxz xz + Cyz yz
                // + D xyz, and its partial derivatives at (x0, y0, z0)
                const float A = cv[0];
                const float Bz = cv[1] - cv[0], e1 = cv[3] - cv[2], e2 = cv[5] - cv[4], e3 = cv[7] - cv[6];
                const float By = cv[2] - cv[0], g1 = cv[6] - cv[4];
                const float Cyz = e1 - Bz, g3 = e3 - e2;
                const float Bx = cv[4] - cv[0], Cxy = g1 - By, Cxz = e2 - Bz, D = g3 - Cyz;
                const float fxy = __builtin_fmaf(D, z0, Cxy);
                const float fxz = __builtin_fmaf(D, y0, Cxz);
                const float fyz = __builtin_fmaf(D, x0, Cyz);
                const float tyz = __builtin_fmaf(Cyz, z0, By);
                const float fx = __builtin_fmaf(y0, fxy, __builtin_fmaf(Cxz, z0, Bx));
                const float fy = __builtin_fmaf(x0, fxy, tyz);
                const float fz = __builtin_fmaf(x0, fxz, __builtin_fmaf(Cyz, y0, Bz));
                const float k0 = __builtin_fmaf(x0, fx, __builtin_fmaf(y0, tyz, __builtin_fmaf(z0, Bz, A)));
                const float q1 = dtf * idb, q2 = q1 * dtf, q3 = q2 * dtf;
                const float ku[4] = {
                    __builtin_fmaf(k0, idb, qc1),
                    __builtin_fmaf(fx, P0, __builtin_fmaf(fy, P1, fz * P2)) * q1,
                    __builtin_fmaf(fxy, P01, __builtin_fmaf(fxz, P02, fyz * P12)) * q2,
                    D * P012 * q3};
                // quad_visit_ok on k0 and the scaled higher coefficients
                const float mm = __builtin_fabsf(k0) + __builtin_fabsf(ku[1]) + __builtin_fabsf(ku[2]) +
                                 __builtin_fabsf(ku[3]);
                if (__builtin_amdgcn_classf(mm, (1 << 7) | (1 << 8))) ctf_visit<NS>(tf, lt, ku, dtf, ns, st);
            }
        }
        bool out;
        if (bm) {
            if (a01) {
                asm volatile("");
                c0 += st0;
                t0 += td0;
                r0 -= sd0;
                out = (unsigned)c0 >= D0;
            } else {
                asm volatile("");
                c1 += st1;
                t1 += td1;
                r1 -= sd1;
                out = (unsigned)c1 >= D1;
            }
        } else {
            asm volatile("");
            c2 += st2;
            t2 += td2;
            r2 -= sd2;
            out = (unsigned)c2 >= D2;
        }
        en = tm;
        if (out || ge) break;
    }
}

// InterpolatedProjectionSampler.sample (image_samplers.pyx:368-402): the
// trilinear value of each field at the visit's n_samples midpoint
// sub-samples, times dt, summed into rgba[j], j < min(3, n_fields); no mask
// test.  The interpolant along the sub-samples is the visit's cubic in s, so
// the sum over s = 0..n-1 is n k0 + S1 k1 + S2 k2 + S3 k3 with the power
// sums S1, S2, S3 of 0..n-1, formed once per visit in FP64.
template <int NFP, bool REC>
struct IProjSampler {
    const float *data;
    BrickReg b;
    double pa[3];
    double pb[3];
    double inv_ns;
    int n_samples;
    double s1, s2, s3;
    int nch;
    double st[3];
    Rec *rec;

    __device__ __forceinline__ void operator()(double en, double ex, int c0, int c1, int c2)
    {
        if (REC) rec->visit(en, ex, c0, c1, c2);
        const int d1 = b.d1, d2 = b.d2;
        if (!((unsigned)c0 < (unsigned)b.d0 && (unsigned)c1 < (unsigned)d1 &&
              (unsigned)c2 < (unsigned)d2))
            return;
        if (REC) rec->samples += n_samples;
        const int voff = vertex_base(b, c0, c1, c2);
        const double dt = (ex - en) * inv_ns;
        const double tt = en + 0.5 * dt;
        float dp[3], ds[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int ci = (i == 0) ? c0 : ((i == 1) ? c1 : c2);
            dp[i] = (float)(__builtin_fma(tt, pa[i], pb[i]) - (double)ci);
            ds[i] = (float)(pa[i] * dt);
        }
        float cv[8][NFP];
        load_corners<NFP>(data + (size_t)voff * NFP, cv);
        const double ns = (double)n_samples;
#pragma unroll
        for (int f = 0; f < (NFP < 3 ? NFP : 3); f++) {
            if (f >= nch) break;
            const float cf[8] = {cv[0][f], cv[1][f], cv[2][f], cv[3][f],
                                 cv[4][f], cv[5][f], cv[6][f], cv[7][f]};
            float k[4];
            trilinear_cubic(cf, dp, ds, k);
            const double sum = __builtin_fma(
                s3, (double)k[3], __builtin_fma(s2, (double)k[2], __builtin_fma(s1, (double)k[1], ns * (double)k[0])));
            st[f] = __builtin_fma(sum, dt, st[f]);
        }
    }
};

// LightSourceRenderSampler.sample (image_samplers.pyx:548-590) with
// FIT_eval_transfer_with_light (field_interpolation_tables.pxd:115-143): per
// sub-sample the tables' values, the gradient of field 0 normalised and
// negated (eval_gradient, fixed_interpolator.cpp:116-148), dot = light . grad,
// and rgba_c = (1 - ta) e_c (1 + dot l_c) + ta rgba_c; no mask test.  The
// reference's one-sided and central differences of the trilinear
// interpolant along an axis are its exact partial derivative along that axis
// (the interpolant is linear in each coordinate), so the gradient is formed
// analytically from the corner differences.  Where the field is flat to
// within rounding, the reference normalises rounding noise into an
// arbitrary direction; there the two light terms differ.
template <int NFP, bool REC>
struct LightSampler {
    const float *data;
    BrickReg b;
    double pa[3];
    double pb[3];
    const TFCache *tf;
    const float4 *lt;
    double inv_ns;
    int n_samples;
    float ldir[3];
    double lrgba[3];
    double st[3];
    Rec *rec;

    __device__ __forceinline__ void operator()(double en, double ex, int c0, int c1, int c2)
    {
        if (REC) rec->visit(en, ex, c0, c1, c2);
        const int d1 = b.d1, d2 = b.d2;
        if (!((unsigned)c0 < (unsigned)b.d0 && (unsigned)c1 < (unsigned)d1 &&
              (unsigned)c2 < (unsigned)d2))
            return;
        if (REC) rec->samples += n_samples;
        const int voff = vertex_base(b, c0, c1, c2);
        const double dt = (ex - en) * inv_ns;
        const double tt = en + 0.5 * dt;
        float dp[3], ds[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int ci = (i == 0) ? c0 : ((i == 1) ? c1 : c2);
            dp[i] = (float)(__builtin_fma(tt, pa[i], pb[i]) - (double)ci);
            ds[i] = (float)(pa[i] * dt);
        }
        float cv[8][NFP];
        load_corners<NFP>(data + (size_t)voff * NFP, cv);
        float k[NFP][4];
#pragma unroll
        for (int f = 0; f < NFP; f++) {
            const float cf[8] = {cv[0][f], cv[1][f], cv[2][f], cv[3][f],
                                 cv[4][f], cv[5][f], cv[6][f], cv[7][f]};
            trilinear_cubic(cf, dp, ds, k[f]);
        }
        // Field 0 as A + Bx x + By y + Bz z + Cxy xy + Cxz xz + Cyz yz + D xyz.
        const float Bz = cv[1][0] - cv[0][0];
        const float By = cv[2][0] - cv[0][0];
        const float Bx = cv[4][0] - cv[0][0];
        const float Cyz = (cv[3][0] - cv[2][0]) - Bz;
        const float Cxz = (cv[5][0] - cv[4][0]) - Bz;
        const float Cxy = (cv[6][0] - cv[4][0]) - By;
        const float D = ((cv[7][0] - cv[6][0]) - (cv[5][0] - cv[4][0])) - Cyz;
        const TFCache &T = *tf;
        for (int s = 0; s < n_samples; s++) {
            const float sf = (float)s;
            float dvs[NFP], e[4];
#pragma unroll
            for (int f = 0; f < NFP; f++)
                dvs[f] = __builtin_fmaf(
                    __builtin_fmaf(__builtin_fmaf(k[f][3], sf, k[f][2]), sf, k[f][1]), sf, k[f][0]);
            const float x = __builtin_fmaf(sf, ds[0], dp[0]);
            const float y = __builtin_fmaf(sf, ds[1], dp[1]);
            const float z = __builtin_fmaf(sf, ds[2], dp[2]);
            const float gx = __builtin_fmaf(D * y, z, __builtin_fmaf(Cxz, z, __builtin_fmaf(Cxy, y, Bx)));
            const float gy = __builtin_fmaf(D * x, z, __builtin_fmaf(Cyz, z, __builtin_fmaf(Cxy, x, By)));
            const float gz = __builtin_fmaf(D * x, y, __builtin_fmaf(Cyz, y, __builtin_fmaf(Cxz, x, Bz)));
            const float n2 = __builtin_fmaf(gx, gx, __builtin_fmaf(gy, gy, gz * gz));
            double dot = 0.0;
            if (n2 != 0.0f) {
                const double lg = (double)__builtin_fmaf(ldir[0], gx, __builtin_fmaf(ldir[1], gy, ldir[2] * gz));
                dot = -lg / __builtin_sqrt((double)n2);
            }
            fit_eval_general<NFP>(T, lt, dvs, e);
            if (T.grey) {
                const double ta = y_fmax(1.0 - dt * ((double)e[0] + (double)e[1] + (double)e[2]), 0.0);
#pragma unroll
                for (int c = 0; c < 3; c++)
                    st[c] = (1.0 - ta) * (double)e[c] * (1.0 + dot * lrgba[c]) + ta * st[c];
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double ta = y_fmax(1.0 - dt * (double)e[c], 0.0);
                    st[c] = (1.0 - ta) * (double)e[c] * (1.0 + dot * lrgba[c]) + ta * st[c];
                }
            }
        }
    }
};

// Casts the rays of one 8x8 tile of image `im` through its bricks.  `im`
// and `tile` are wave-uniform.
template <int KIND, int NFP, bool REC, int MODE>
__device__ __forceinline__ void render_tile(const DImg &im, int tile, int lane,
                                            const float4 *lt, float *wbuf = nullptr, int wave_words = 0)
{
    const int tsh = im.tsh;
    const int tw = 1 << tsh, th = 64 >> tsh;
    const int tx = tile / im.tiles1;
    const int ty = tile - tx * im.tiles1;
    const int ti0 = tx * th, tj0 = ty * tw;
    const int vi = ti0 + (lane >> tsh);
    const int vj = tj0 + (lane & (tw - 1));
    const bool active = (vi < im.nv0) && (vj < im.nv1);
    const int64_t npix = (int64_t)im.nv0 * im.nv1;
    const int64_t pix = active ? (int64_t)vi * im.nv1 + vj : 0;

    // MODE bit 3: per-pixel lens (generate_vector_info_null,
    // lenses.pyx:196-207); rays no longer share a direction.
    constexpr bool PP = (MODE & 8) != 0;
    double pos[3], dir[3], iv[3];
    if constexpr (PP) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            pos[i] = im.pix_pos[pix * 3 + i];
            dir[i] = im.pix_dir[pix * 3 + i];
        }
    } else {
        ray_setup(im, vi, vj, pos, dir);
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double v = (dir[i] != 0.0) ? 1.0 / dir[i] : 0.0;
        iv[i] = PP ? v : uni_d(v);
    }
    double max_t = 1.0;
    if (im.zbuf != nullptr && active) max_t = y_fmin(y_fmax(im.zbuf[pix], 0.0), 1.0);

    Rec rec;
    rec.visits = 0;
    rec.samples = 0;
    rec.hash = 0;
    rec.brick = 0;

    TFCache tfc;
    int nb = 0, tform = 0;
    if (KIND == 1 || KIND == 3) {
        load_tf(*im.tf, tfc);
        nb = uni(im.tf->nbins[0]);
        tform = uni(im.tform);
    }

    float ps[NFP], pc[NFP];
    double st[4];
#pragma unroll
    for (int f = 0; f < NFP; f++) {
        ps[f] = 0.0f;
        pc[f] = 0.0f;
    }
#pragma unroll
    for (int c = 0; c < 4; c++) st[c] = 0.0;
    if (im.init != nullptr && active) {
        if (KIND == 0) {
#pragma unroll
            for (int f = 0; f < NFP; f++)
                if (f < 4) ps[f] = im.init[f * npix + pix];
        } else {
#pragma unroll
            for (int c = 0; c < 4; c++) st[c] = (double)im.init[c * npix + pix];
        }
    }
    if (KIND == 1 && tform) {
#pragma unroll
        for (int c = 0; c < 3; c++) st[c] = 1.0 - st[c];
    }

    // Early ray termination (YTRC_FLAG_ERT, VR only): the reference composites
    // each new sample on top as t grows (field_interpolation_tables.pxd:107-110),
    // so its eye is at t -> +inf.  Front to back is therefore the reverse brick
    // order; each brick's cells are still walked forward (the same DDA) to form
    // its per-channel affine map rgba -> C_b + T_b rgba, maps are composed as
    // C += T C_b, T *= T_b, and a wave stops when every lane has T < 1e-4.
    // A partial-map request (YTRC_FLAG_PARTIAL) composes the same maps over
    // all of its bricks, without termination unless YTRC_FLAG_ERT is also
    // set, and stores (C, T) instead of C + T rgba_in.
    constexpr bool ert = (KIND == 1) && ((MODE & 4) != 0);
    const double ert_eps = (im.flags & YTRC_FLAG_ERT) ? 1e-4 : -1.0;
    double rgba0[4], Cacc[4], Tacc[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        rgba0[c] = st[c];
        Cacc[c] = 0.0;
        Tacc[c] = 1.0;
    }
    if (ert && tform) {
#pragma unroll
        for (int c = 0; c < 3; c++) rgba0[c] = 1.0 - st[c];
    }

    for (int kk = 0; kk < im.n_order; kk++) {
        const int k = ert ? im.n_order - 1 - kk : kk;
        const int4 e = im.ext[k];
        if (e.y <= ti0 || e.x >= ti0 + th || e.w <= tj0 || e.z >= tj0 + tw) continue;
        bool go = active && vi >= e.x && vi < e.y && vj >= e.z && vj < e.w;
        const int bid = im.order[k];
        const DBrick &b = im.bricks[bid];
        Walk w;
        if (go) go = walk_init<!PP>(b, pos, dir, iv, max_t, w);
        if (go) {
            rec.brick = bid;
            if (KIND == 0) {
                ProjSampler<NFP, REC> s;
                s.data = im.pool + b.data_off;
                s.d1 = uni(b.dims[1]);
                s.d2 = uni(b.dims[2]);
                s.dims0 = uni(b.dims[0]);
                s.nbj = (s.d1 + 3) >> 2;
                s.nbk = (s.d2 + 3) >> 2;
#pragma unroll
                for (int f = 0; f < NFP; f++) {
                    s.s[f] = ps[f];
                    s.c[f] = pc[f];
                }
                s.rec = &rec;
                bool fastw = false;
                if constexpr (!REC && NFP <= 4) {
                    // the fast walk's conditions (walk_projection_fast)
                    const unsigned nb0 = ((unsigned)s.dims0 + 3u) >> 2;
                    const uint64_t bytes = (uint64_t)nb0 * (unsigned)s.nbj * (unsigned)s.nbk * 64u * 4u * NFP;
                    const uint64_t k0 = (uint64_t)s.nbj * (unsigned)s.nbk * 64u;
                    const bool inside = (unsigned)w.cur[0] < (unsigned)s.dims0 &&
                                        (unsigned)w.cur[1] < (unsigned)s.d1 &&
                                        (unsigned)w.cur[2] < (unsigned)s.d2;
                    if (bytes < (1ull << 32) && k0 < (1u << 24) && __all(inside)) {
                        const __amdgpu_buffer_rsrc_t rs = brick_rsrc(s.data, (unsigned)bytes);
                        const unsigned K0 = (unsigned)k0 - 64u, K1 = (unsigned)s.nbk * 64u - 16u;
                        if (wbuf != nullptr && wave_words >= PTAB_WAVE && s.dims0 <= PTAB && s.d1 <= PTAB &&
                            s.d2 <= PTAB)
                            walk_projection_fast<NFP, true>(w, max_t, s, rs, K0, K1, (unsigned *)wbuf);
                        else
                            walk_projection_fast<NFP>(w, max_t, s, rs, K0, K1);
                        fastw = true;
                    }
                }
                if (!fastw) walk_projection(b, w, max_t, s);
#pragma unroll
                for (int f = 0; f < NFP; f++) {
                    ps[f] = s.s[f];
                    pc[f] = s.c[f];
                }
            } else if constexpr (KIND == 3) {
                LightSampler<NFP, REC> s;
                s.data = im.pool + b.data_off;
                s.b = brick_reg(b);
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    s.pa[i] = PP ? dir[i] * b.idds[i] : uni_d(dir[i] * b.idds[i]);
                    s.pb[i] = (pos[i] - b.le[i]) * b.idds[i];
                    s.ldir[i] = (float)im.ldir[i];
                    s.lrgba[i] = im.lrgba[i];
                }
                s.tf = &tfc;
                s.lt = lt;
                s.inv_ns = im.inv_ns;
                s.n_samples = im.n_samples;
#pragma unroll
                for (int c = 0; c < 3; c++) s.st[c] = st[c];
                s.rec = &rec;
                walk_loop(b, w, max_t, s);
#pragma unroll
                for (int c = 0; c < 3; c++) st[c] = s.st[c];
            } else if constexpr (KIND == 2) {
                IProjSampler<NFP, REC> s;
                s.data = im.pool + b.data_off;
                s.b = brick_reg(b);
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    s.pa[i] = PP ? dir[i] * b.idds[i] : uni_d(dir[i] * b.idds[i]);
                    s.pb[i] = (pos[i] - b.le[i]) * b.idds[i];
                }
                s.inv_ns = im.inv_ns;
                s.n_samples = im.n_samples;
                const double ns = (double)im.n_samples;
                s.s1 = 0.5 * ns * (ns - 1.0);
                s.s2 = (ns - 1.0) * ns * (2.0 * ns - 1.0) / 6.0;
                s.s3 = s.s1 * s.s1;
                s.nch = im.nch_out;
#pragma unroll
                for (int c = 0; c < 3; c++) s.st[c] = st[c];
                s.rec = &rec;
                walk_loop(b, w, max_t, s);
#pragma unroll
                for (int c = 0; c < 3; c++) st[c] = s.st[c];
            } else {
                VRSampler<NFP, REC, MODE> s;
                s.data = im.pool + b.data_off;
                s.mask = (b.mask_off >= 0) ? im.masks + b.mask_off : nullptr;
                s.b = brick_reg(b);
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    s.pa[i] = PP ? dir[i] * b.idds[i] : uni_d(dir[i] * b.idds[i]);
                    s.pb[i] = (pos[i] - b.le[i]) * b.idds[i];
                    s.paf[i] = (float)s.pa[i];
                }
                s.tf = &tfc;
                s.lt = lt;
                s.nb = nb;
                s.tform = tform;
                s.inv_ns = im.inv_ns;
                s.n_samples = im.n_samples;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    s.st[c] = ert ? (tform ? 1.0 : 0.0) : st[c];
                    s.tb[c] = 1.0;
                }
                s.rec = &rec;
                bool fastw = false;
                if constexpr (!REC && NFP == 1 && (MODE & 1) != 0 && !PP) {
                    // the fast walk's conditions (walk_vr_ctf_fast)
                    const uint64_t bytes = (uint64_t)vblocks(s.b.d0) * (unsigned)s.b.vnbj *
                                           (unsigned)s.b.vnbk * (VBP * 4u);
                    const uint64_t k0 = vk0_bytes((unsigned)s.b.vnbj, (unsigned)s.b.vnbk);
                    const bool inside = (unsigned)w.cur[0] < (unsigned)s.b.d0 &&
                                        (unsigned)w.cur[1] < (unsigned)s.b.d1 &&
                                        (unsigned)w.cur[2] < (unsigned)s.b.d2;
                    const uint64_t k1 = vk1_bytes((unsigned)s.b.vnbk);
                    if (bytes < (1ull << 32) && k0 < (1u << 24) && k1 < (1u << 24) && __all(inside)) {
                        walk_vr_ctf_fast<(MODE & 2) ? 10 : 0, (MODE & 16) != 0>(
                            w, max_t, s.b, brick_rsrc(s.data, (unsigned)bytes), (unsigned)k0,
                            (unsigned)k1, s.mask, s.pa, s.pb, tfc, lt,
                            s.inv_ns, s.n_samples, s.st, wbuf);
                        fastw = true;
                    }
                }
                if (!fastw) walk_loop(b, w, max_t, s);
                if constexpr (ert) {
                    double cb[4], tb[4];
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        if (tform) {
                            cb[c] = (c < 3) ? 1.0 - s.st[c] : 0.0;
                            tb[c] = (c < 3) ? s.st[c] : 1.0;
                        } else if (tfc.grey) {
                            cb[c] = s.st[c];
                            tb[c] = s.tb[0];
                        } else {
                            cb[c] = (c < 3) ? s.st[c] : 0.0;
                            tb[c] = (c < 3) ? s.tb[c] : 1.0;
                        }
                        Cacc[c] = __builtin_fma(Tacc[c], cb[c], Cacc[c]);
                        Tacc[c] *= tb[c];
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < 4; c++) st[c] = s.st[c];
                }
            }
        }
        if constexpr (ert) {
            const double tm = y_fmax(y_fmax(Tacc[0], Tacc[1]), y_fmax(Tacc[2], tfc.grey ? Tacc[3] : 0.0));
            if (__all(!active || tm < ert_eps)) break;
        }
    }
    if (!active) return;
    if (KIND == 0) {
#pragma unroll
        for (int f = 0; f < NFP; f++)
            if (f < im.nch_out) im.out[f * npix + pix] = ps[f] - pc[f];
    } else if constexpr (ert) {
        if (im.flags & YTRC_FLAG_PARTIAL) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                im.out[c * npix + pix] = (float)Cacc[c];
                im.out[(4 + c) * npix + pix] = (float)Tacc[c];
            }
        } else {
#pragma unroll
            for (int c = 0; c < 4; c++)
                if (c < im.nch_out) im.out[c * npix + pix] = (float)__builtin_fma(Tacc[c], rgba0[c], Cacc[c]);
        }
    } else {
        if (tform) {
#pragma unroll
            for (int c = 0; c < 3; c++) st[c] = 1.0 - st[c];
        }
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (c < im.nch_out) im.out[c * npix + pix] = (float)st[c];
    }
    if (REC) {
        im.rec_visits[pix] = rec.visits;
        im.rec_hash[pix] = rec.hash;
        im.rec_samples[pix] = rec.samples;
    }
}

// Plain launch: one request; blockDim.x/64 tiles per workgroup.
template <int KIND, int NFP, bool REC, int MODE>
__global__ __launch_bounds__(256) void k_render(const DImg *__restrict__ imgs, int idx)
{
    extern __shared__ float4 s_tf[];
    const DImg &im = imgs[idx];
    if (KIND == 1 || KIND == 3) {
        const DTF &T = *im.tf;
        for (int i = threadIdx.x; i < T.total_bins; i += blockDim.x) s_tf[i] = T.table[i];
        __syncthreads();
    }
    const int waves = blockDim.x >> 6;
    const int tile = blockIdx.x * waves + (threadIdx.x >> 6);
    if (tile >= im.tiles0 * im.tiles1) return;
    render_tile<KIND, NFP, REC, MODE>(im, tile, threadIdx.x & 63, s_tf);
}

// Tile-queue partitions of a persistent launch: one per XCD of the MI355X
// (256 CUs in 8 XCDs of 32, each XCD with its own L2).
constexpr int QSPLIT = 8;
// A launch's QSPLIT queue counters sit 256 bytes apart, each on a line of
// its own, so that the atomics of the eight ranges do not queue on one line.
constexpr int QSTRIDE = 64;
constexpr int QCOUNTERS = QSPLIT * QSTRIDE;  // words of counters per launch

// Persistent launch: every wave repeatedly takes the next (image, tile)
// item of a global queue until the queue is drained.  The launch's
// transfer-function tables (tf_all, tf_bins: those of its requests) sit in
// LDS, each image at its own slot, followed by wave_words words per wave
// (the vertex-block buffer with MODE bit 4, or a walk's offset table).  The item
// index is made wave-uniform so that image, brick and table descriptors are
// read with scalar loads.
// Minimum waves per SIMD requested from the register allocator, 6 (80
// VGPRs) for the production variants:
// - volume render, one field, ColorTransferFunction path, no record/ERT
//   (with or without vertex blocks in LDS): spills outside the visit loop;
//   with the fast walk 2.3 % faster than 5 waves and within 0.5 % of 7;
//   before it, 0.8 % faster than 5 waves (95 VGPRs, no spill), 2.5 %
//   faster than 4, and 7 waves (72 VGPRs, 68-byte spill) 0.5 % slower;
// - projection, one field, no record: 3.6 % faster than the 5 waves its
//   90 VGPRs give.
template <int KIND, int NFP, bool REC, int MODE>
struct MinWaves {
    static constexpr int value = (KIND == 1 && NFP == 1 && !REC && (MODE == 1 || MODE == 3 || MODE == 19)) ? 6
                                 : (KIND == 0 && NFP == 1 && !REC)                          ? 6
                                                                                             : 1;
};

template <int KIND, int NFP, bool REC, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MinWaves<KIND, NFP, REC, MODE>::value)))
void k_persist(const DImg *__restrict__ imgs,
                                                 const int2 *__restrict__ items, int n_items,
                                                 unsigned int *__restrict__ counter,
                                                 const float4 *__restrict__ tf_all, int tf_bins, int wave_words)
{
    extern __shared__ float4 s_tf[];
    if (KIND == 1 || KIND == 3) {
        for (int i = threadIdx.x; i < tf_bins; i += blockDim.x) s_tf[i] = tf_all[i];
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    // This wave's wave_words words of LDS after the tables (none if 0): with
    // MODE bit 4 its 512-byte vertex-block buffer, for a projection its table
    // of offset terms (walk_projection_fast uses it if it is large enough).
    float *wbuf = (wave_words > 0) ? (float *)(s_tf + tf_bins) + (threadIdx.x >> 6) * wave_words : nullptr;
    // The item list is split into QSPLIT contiguous ranges with a counter
    // each.  Workgroups are dispatched round-robin over the XCDs, so
    // blockIdx.x % QSPLIT names this workgroup's XCD: it drains the range of
    // neighbouring tiles that XCD's L2 already holds lines for, then helps
    // with the other ranges in turn.  Every wave exits after finding all
    // QSPLIT ranges empty.
    int part = (int)(blockIdx.x % QSPLIT);
    int left = QSPLIT;
    // Items this wave has taken, started at a per-wave phase so that the
    // waves' reads of the stop flag are spread over time (below).
    int taken = (int)((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & 15);
    while (true) {
        unsigned int it = 0;
        if (lane == 0) it = atomicAdd(counter + part * QSTRIDE, 1u);
        it = (unsigned int)__builtin_amdgcn_readfirstlane((int)__shfl(it, 0));
        const unsigned int lo = (unsigned int)(((int64_t)n_items * part) / QSPLIT);
        const unsigned int hi = (unsigned int)(((int64_t)n_items * (part + 1)) / QSPLIT);
        if (it >= hi - lo) {
            if (--left == 0) break;
            part = (part + 1 == QSPLIT) ? 0 : part + 1;
            continue;
        }
        const int2 w = items[lo + it];
        const int img = __builtin_amdgcn_readfirstlane(w.x);
        const int tword = __builtin_amdgcn_readfirstlane(w.y);
        // an item holds 1 tile (count field 0 or 1) or a run of them
        const int tile = tword & 0xffffff;
        const int cnt = max(1, tword >> 24);
        const DImg &im = imgs[img];
        // A stop request ends the wave at one of every 16th item it takes,
        // each wave at its own phase.  The flag sits in host memory: read
        // per item by every wave, or by every wave at its first item, those
        // reads queue on the PCIe link (the projection kernel took 2x and
        // 1.1x its time).
        if ((taken++ & 15) == 15 && uni((int)stop_requested(im.stop))) break;
        for (int t = 0; t < cnt; t++) render_tile<KIND, NFP, REC, MODE>(im, tile + t, lane, s_tf + im.tf_slot, wbuf, wave_words);
    }
}

}  // namespace
