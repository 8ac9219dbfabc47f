/*
 * ytrc_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 restatement of yt's brick ray-caster, used as the parity
 * checker for the HIP product path and as the CPU baseline timed by bench.py.
 * It is never linked into, loaded by, or called from the product library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp, no fast-math),
 * which reproduces the reference's strict IEEE FP64 behaviour on x86-64
 * (no FMA contraction: yt is built without -march, setup.py:71-87).
 *
 * Every function cites the reference file:line it restates.  Parity of each
 * function is pinned by the tests listed in its header comment.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* libm's M_PI (math.h outside strict ISO C) */
#define ORC_PI 3.14159265358979323846

/* ---------------------------------------------------------------------- */
/* fp_utils.pxd:17-45 -- yt's own fmin/fmax/fclip/iclip (non-IEEE NaN rules) */
static inline double y_fmax(double f0, double f1) { return (f0 > f1) ? f0 : f1; }
static inline double y_fmin(double f0, double f1) { return (f0 < f1) ? f0 : f1; }
static inline double y_fclip(double f, double a, double b) { return y_fmin(y_fmax(f, a), b); }
static inline int64_t y_iclip(int64_t i, int64_t a, int64_t b)
{
    if (i < a) return a;
    if (i > b) return b;
    return i;
}

/* volume_container.pxd:13-33 -- the VolumeContainer brick descriptor. */
typedef struct {
    int n_fields;
    const double *const *data; /* n_fields pointers, C-order arrays        */
    const uint8_t *mask;       /* dims-shaped uint8, or NULL               */
    double left_edge[3];
    double right_edge[3];
    double dds[3];
    double idds[3];
    int dims[3];
} orc_vc;

/* partitioned_grid.pyx:49-55: dds=(RE-LE)/dims, idds=1/dds. */
static void orc_vc_init(orc_vc *c, const double *le, const double *re, const int32_t *dims)
{
    for (int i = 0; i < 3; i++) {
        c->left_edge[i] = le[i];
        c->right_edge[i] = re[i];
        c->dims[i] = dims[i];
        c->dds[i] = (c->right_edge[i] - c->left_edge[i]) / (double)dims[i];
        c->idds[i] = 1.0 / c->dds[i];
    }
}

/* grid_traversal.pxd:18-25 -- sampler callback signature. */
typedef void (*orc_sampler)(const orc_vc *vc, const double v_pos[3], const double v_dir[3],
                            double enter_t, double exit_t, const int index[3], void *data);

/*
 * walk_volume: grid_traversal.pyx:26-140 (Amanatides & Woo DDA).
 * Pinned by tests/test_oracle_walk.py: exact-rational slab brute force on
 * tiny grids (the test_rays.py:42-59 method), sum(dt) == clipped chord,
 * monotone t, unit steps, and the tie-breaking cases.
 * Deviation: iv_dir is zero-initialised (the reference leaves it unset for
 * axes with v_dir == 0, where it is never used for a selected axis).
 */
static int orc_walk_volume(const orc_vc *vc, const double v_pos[3], const double v_dir[3],
                           orc_sampler sample, void *data, double max_t)
{
    int cur_ind[3], step[3];
    int x, y, i, hit, direction;
    double intersect_t = 1.1;
    double iv_dir[3] = {0.0, 0.0, 0.0};
    double tmax[3], tdelta[3];
    double exit_t = -1.0, enter_t = -1.0;
    double tl, temp_x, temp_y = -1;
    if (max_t > 1.0) max_t = 1.0;
    direction = -1;
    if (vc->left_edge[0] <= v_pos[0] && v_pos[0] < vc->right_edge[0] &&
        vc->left_edge[1] <= v_pos[1] && v_pos[1] < vc->right_edge[1] &&
        vc->left_edge[2] <= v_pos[2] && v_pos[2] < vc->right_edge[2]) {
        intersect_t = 0.0;
        direction = 3;
    }
    for (i = 0; i < 3; i++) {
        if (v_dir[i] < 0) {
            step[i] = -1;
        } else if (v_dir[i] == 0.0) {
            step[i] = 0;
            continue;
        } else {
            step[i] = 1;
        }
        iv_dir[i] = 1.0 / v_dir[i];
        if (direction == 3) continue;
        x = (i + 1) % 3;
        y = (i + 2) % 3;
        if (step[i] > 0)
            tl = (vc->left_edge[i] - v_pos[i]) * iv_dir[i];
        else
            tl = (vc->right_edge[i] - v_pos[i]) * iv_dir[i];
        temp_x = (v_pos[x] + tl * v_dir[x]);
        temp_y = (v_pos[y] + tl * v_dir[y]);
        if (fabs(temp_x - vc->left_edge[x]) < 1e-10 * vc->dds[x])
            temp_x = vc->left_edge[x];
        else if (fabs(temp_x - vc->right_edge[x]) < 1e-10 * vc->dds[x])
            temp_x = vc->right_edge[x];
        if (fabs(temp_y - vc->left_edge[y]) < 1e-10 * vc->dds[y])
            temp_y = vc->left_edge[y];
        else if (fabs(temp_y - vc->right_edge[y]) < 1e-10 * vc->dds[y])
            temp_y = vc->right_edge[y];
        if (vc->left_edge[x] <= temp_x && temp_x <= vc->right_edge[x] &&
            vc->left_edge[y] <= temp_y && temp_y <= vc->right_edge[y] &&
            0.0 <= tl && tl < intersect_t) {
            direction = i;
            intersect_t = tl;
        }
    }
    if (enter_t >= 0.0) intersect_t = enter_t;
    if (!((0.0 <= intersect_t) && (intersect_t < max_t))) return 0;
    for (i = 0; i < 3; i++) {
        tdelta[i] = (double)step[i] * iv_dir[i] * vc->dds[i];
        if (i == direction && step[i] > 0) {
            cur_ind[i] = 0;
        } else if (i == direction && step[i] < 0) {
            cur_ind[i] = vc->dims[i] - 1;
        } else {
            temp_x = intersect_t * v_dir[i] + v_pos[i];
            temp_y = ((temp_x - vc->left_edge[i]) * vc->idds[i]);
            if (-1 < temp_y && temp_y < 0 && step[i] > 0)
                temp_y = 0.0;
            else if ((double)(vc->dims[i] - 1) < temp_y && temp_y < (double)vc->dims[i] && step[i] < 0)
                temp_y = (double)(vc->dims[i] - 1);
            cur_ind[i] = (int)(floor(temp_y));
        }
        if (step[i] > 0)
            temp_y = (double)(cur_ind[i] + 1) * vc->dds[i] + vc->left_edge[i];
        else if (step[i] < 0)
            temp_y = (double)cur_ind[i] * vc->dds[i] + vc->left_edge[i];
        tmax[i] = (temp_y - v_pos[i]) * iv_dir[i];
        if (step[i] == 0) tmax[i] = 1e60;
    }
    for (i = 0; i < 3; i++) {
        if (cur_ind[i] == vc->dims[i] && step[i] >= 0) return 0;
        if (cur_ind[i] == -1 && step[i] <= -1) return 0;
    }
    enter_t = intersect_t;
    hit = 0;
    while (1) {
        hit += 1;
        if (tmax[0] < tmax[1]) {
            i = (tmax[0] < tmax[2]) ? 0 : 2;
        } else {
            i = (tmax[1] < tmax[2]) ? 1 : 2;
        }
        exit_t = y_fmin(tmax[i], max_t);
        sample(vc, v_pos, v_dir, enter_t, exit_t, cur_ind, data);
        cur_ind[i] += step[i];
        enter_t = tmax[i];
        tmax[i] += tdelta[i];
        if (cur_ind[i] < 0 || cur_ind[i] >= vc->dims[i] || enter_t >= max_t) break;
    }
    return hit;
}

/* ---------------------------------------------------------------------- */
/* FieldInterpolationTable: field_interpolation_tables.pxd:27-61 */
typedef struct {
    const double *values;
    double bounds[2];
    double dbin;
    double idbin;
    double *d0;
    double *dy;
    int field_id;
    int weight_field_id;
    int weight_table_id;
    int nbins;
} orc_fit;

/* FIT_initialize_table: field_interpolation_tables.pxd:44-61.
 * Pinned by tests/test_oracle_fit.py (exact at nodes, linear between). */
static void orc_fit_init(orc_fit *fit, int nbins, const double *values, double b1, double b2,
                         int field_id, int weight_field_id, int weight_table_id)
{
    fit->bounds[0] = b1;
    fit->bounds[1] = b2;
    fit->nbins = nbins;
    fit->dbin = (fit->bounds[1] - fit->bounds[0]) / (double)(fit->nbins - 1);
    fit->idbin = 1.0 / fit->dbin;
    fit->values = values;
    fit->d0 = (double *)calloc((size_t)nbins, sizeof(double));
    fit->dy = (double *)calloc((size_t)nbins, sizeof(double));
    for (int i = 0; i < nbins - 1; i++) {
        fit->d0[i] = fit->bounds[0] + (double)i * fit->dbin;
        fit->dy[i] = (fit->values[i + 1] - fit->values[i]) * fit->idbin;
    }
    fit->field_id = field_id;
    fit->weight_field_id = weight_field_id;
    fit->weight_table_id = weight_table_id;
}

/* FIT_get_value: field_interpolation_tables.pxd:66-80.
 * 0 outside the open interval (b0,b1) and for non-normal inputs
 * (platform_dep_math.hpp:14-19 -> std::isnormal). */
static inline double orc_fit_get_value(const orc_fit *fit, const double dvs[6])
{
    double dd, dout;
    int bin_id;
    double v = dvs[fit->field_id];
    if (v >= fit->bounds[1] || v <= fit->bounds[0]) return 0.0;
    if (!isnormal(v)) return 0.0;
    bin_id = (int)((v - fit->bounds[0]) * fit->idbin);
    bin_id = (int)y_iclip(bin_id, 0, fit->nbins - 2);
    dd = v - fit->d0[bin_id];
    dout = fit->values[bin_id] + dd * fit->dy[bin_id];
    int wfi = fit->weight_field_id;
    if (wfi != -1) dout *= dvs[wfi];
    return dout;
}

/* FIT_eval_transfer: field_interpolation_tables.pxd:85-110. */
static inline void orc_fit_eval_transfer(double dt, const double *dvs, double *rgba, int n_fits,
                                         const orc_fit *fits, const int field_table_ids[6],
                                         int grey_opacity)
{
    int i, fid;
    double ta;
    double istorage[6] = {0, 0, 0, 0, 0, 0};
    double trgba[6];
    for (i = 0; i < n_fits; i++) istorage[i] = orc_fit_get_value(&fits[i], dvs);
    for (i = 0; i < n_fits; i++) {
        fid = fits[i].weight_table_id;
        if (fid != -1) istorage[i] *= istorage[fid];
    }
    for (i = 0; i < 6; i++) trgba[i] = istorage[field_table_ids[i]];
    if (grey_opacity == 1) {
        ta = y_fmax(1.0 - dt * trgba[3], 0.0);
        for (i = 0; i < 4; i++) rgba[i] = dt * trgba[i] + ta * rgba[i];
    } else {
        for (i = 0; i < 3; i++) {
            ta = y_fmax(1.0 - dt * trgba[i], 0.0);
            rgba[i] = dt * trgba[i] + ta * rgba[i];
        }
    }
}

/* offset_interpolate: fixed_interpolator.cpp:34-52 (OINDEX :13), lerp z->y->x.
 * Pinned by tests/test_oracle_fit.py (exact on linear fields, corners). */
static inline double orc_offset_interpolate(const int ds[3], const double dp[3], const double *data)
{
    const int s1 = (ds[2] + 1);
    const int s0 = (ds[1] + 1) * (ds[2] + 1);
#define OIDX(A, B, C) data[(A) * s0 + (B) * s1 + (C)]
    double dv, vz[4];
    dv = 1.0 - dp[2];
    vz[0] = dv * OIDX(0, 0, 0) + dp[2] * OIDX(0, 0, 1);
    vz[1] = dv * OIDX(0, 1, 0) + dp[2] * OIDX(0, 1, 1);
    vz[2] = dv * OIDX(1, 0, 0) + dp[2] * OIDX(1, 0, 1);
    vz[3] = dv * OIDX(1, 1, 0) + dp[2] * OIDX(1, 1, 1);
    dv = 1.0 - dp[1];
    vz[0] = dv * vz[0] + dp[1] * vz[1];
    vz[1] = dv * vz[2] + dp[1] * vz[3];
    dv = 1.0 - dp[0];
    vz[0] = dv * vz[0] + dp[0] * vz[1];
#undef OIDX
    return vz[0];
}

/* eval_gradient: fixed_interpolator.cpp:116-148.  Finite differences of
 * offset_interpolate at +-0.05 cell widths (one-sided within 0.05 of a
 * face), normalised and negated; 0 where the difference vector is 0.
 * Pinned bit-for-bit against the reference's own compiled function
 * (tests/test_oracle_ref.py). */
static inline void orc_eval_gradient(const int ds[3], const double dp_in[3], const double *data,
                                     double grad[3])
{
    double dp[3] = {dp_in[0], dp_in[1], dp_in[2]};
    double denom, plus, minus, backup, normval = 0.0;
    for (int i = 0; i < 3; i++) {
        backup = dp[i];
        grad[i] = 0.0;
        if (dp[i] >= 0.95) {
            plus = dp[i];
            minus = dp[i] - 0.05;
        } else if (dp[i] <= 0.05) {
            plus = dp[i] + 0.05;
            minus = 0.0;
        } else {
            plus = dp[i] + 0.05;
            minus = dp[i] - 0.05;
        }
        denom = plus - minus;
        dp[i] = plus;
        grad[i] += orc_offset_interpolate(ds, dp, data) / denom;
        dp[i] = minus;
        grad[i] -= orc_offset_interpolate(ds, dp, data) / denom;
        dp[i] = backup;
        normval += grad[i] * grad[i];
    }
    if (normval != 0.0) {
        normval = sqrt(normval);
        for (int i = 0; i < 3; i++) grad[i] /= -normval;
    } else {
        grad[0] = grad[1] = grad[2] = 0.0;
    }
}

/* FIT_eval_transfer_with_light: field_interpolation_tables.pxd:115-143. */
static inline void orc_fit_eval_transfer_with_light(double dt, const double *dvs, const double *grad,
                                                    const double *l_dir, const double *l_rgba,
                                                    double *rgba, int n_fits, const orc_fit *fits,
                                                    const int field_table_ids[6], int grey_opacity)
{
    int i, fid;
    double ta, dot_prod = 0.0;
    double istorage[6] = {0, 0, 0, 0, 0, 0};
    double trgba[6];
    for (i = 0; i < 3; i++) dot_prod += l_dir[i] * grad[i];
    for (i = 0; i < n_fits; i++) istorage[i] = orc_fit_get_value(&fits[i], dvs);
    for (i = 0; i < n_fits; i++) {
        fid = fits[i].weight_table_id;
        if (fid != -1) istorage[i] *= istorage[fid];
    }
    for (i = 0; i < 6; i++) trgba[i] = istorage[field_table_ids[i]];
    if (grey_opacity == 1) {
        ta = y_fmax(1.0 - dt * (trgba[0] + trgba[1] + trgba[2]), 0.0);
        for (i = 0; i < 3; i++)
            rgba[i] = (1. - ta) * trgba[i] * (1. + dot_prod * l_rgba[i]) + ta * rgba[i];
    } else {
        for (i = 0; i < 3; i++) {
            ta = y_fmax(1.0 - dt * trgba[i], 0.0);
            rgba[i] = (1. - ta) * trgba[i] * (1. + dot_prod * l_rgba[i]) + ta * rgba[i];
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Accumulator: image_samplers.pxd:34-36 (ImageAccumulator) plus the
 * VolumeRenderAccumulator of image_samplers.pyx:45-55, and optional
 * per-ray visit instrumentation (count + order-sensitive hash). */
typedef struct {
    int n_fits;
    int n_samples;
    const orc_fit *fits;
    int field_table_ids[6];
    int grey_opacity;
    double light_dir[3];  /* LightSourceRenderSampler: normalised direction */
    double light_rgba[4];
} orc_vra;

typedef struct {
    double rgba[4];
    const orc_vra *vra;
    int brick_id;
    int64_t visits;
    uint64_t hash;
    int64_t samples;
} orc_acc;

static inline int orc_in_range(const orc_vc *vc, const int index[3])
{
    return index[0] >= 0 && index[0] < vc->dims[0] && index[1] >= 0 && index[1] < vc->dims[1] &&
           index[2] >= 0 && index[2] < vc->dims[2];
}

/* ProjectionSampler.sample: image_samplers.pyx:330-343 (no mask check).
 * Deviation: a visit whose index lies outside the brick (possible only for
 * rays within 1e-10*dds outside a face, where the reference reads out of
 * bounds) contributes 0. */
static void orc_projection_sample(const orc_vc *vc, const double v_pos[3], const double v_dir[3],
                                  double enter_t, double exit_t, const int index[3], void *data)
{
    (void)v_pos;
    (void)v_dir;
    orc_acc *im = (orc_acc *)data;
    double dl = (exit_t - enter_t);
    if (!orc_in_range(vc, index)) return;
    int di = (index[0] * vc->dims[1] + index[1]) * vc->dims[2] + index[2];
    int n = vc->n_fields < 4 ? vc->n_fields : 4;
    for (int i = 0; i < n; i++) im->rgba[i] += vc->data[i][di] * dl;
}

/* VolumeRenderSampler.sample: image_samplers.pyx:453-487. */
static void orc_volume_render_sample(const orc_vc *vc, const double v_pos[3], const double v_dir[3],
                                     double enter_t, double exit_t, const int index[3], void *data)
{
    orc_acc *im = (orc_acc *)data;
    const orc_vra *vri = im->vra;
    if (!orc_in_range(vc, index)) return;
    int offset = index[0] * (vc->dims[1] + 1) * (vc->dims[2] + 1) + index[1] * (vc->dims[2] + 1) +
                 index[2];
    int cell_offset = index[0] * (vc->dims[1]) * (vc->dims[2]) + index[1] * (vc->dims[2]) + index[2];
    if (vc->mask != NULL && vc->mask[cell_offset] != 1) return;
    double dp[3], ds[3];
    double dt = (exit_t - enter_t) / (double)vri->n_samples;
    double dvs[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3; i++) {
        dp[i] = (enter_t + 0.5 * dt) * v_dir[i] + v_pos[i];
        dp[i] -= (double)index[i] * vc->dds[i] + vc->left_edge[i];
        dp[i] *= vc->idds[i];
        ds[i] = v_dir[i] * vc->idds[i] * dt;
    }
    for (int s = 0; s < vri->n_samples; s++) {
        for (int j = 0; j < vc->n_fields; j++)
            dvs[j] = orc_offset_interpolate(vc->dims, dp, vc->data[j] + offset);
        orc_fit_eval_transfer(dt, dvs, im->rgba, vri->n_fits, vri->fits, vri->field_table_ids,
                              vri->grey_opacity);
        for (int j = 0; j < 3; j++) dp[j] += ds[j];
    }
}

/* InterpolatedProjectionSampler.sample: image_samplers.pyx:368-402.
 * Vertex-centred data, no mask test; per sub-sample the trilinear value of
 * each field times dt is added to rgba[j], j < min(3, n_fields). */
static void orc_interpolated_projection_sample(const orc_vc *vc, const double v_pos[3],
                                               const double v_dir[3], double enter_t, double exit_t,
                                               const int index[3], void *data)
{
    orc_acc *im = (orc_acc *)data;
    const orc_vra *vri = im->vra;
    if (!orc_in_range(vc, index)) return;
    int offset = index[0] * (vc->dims[1] + 1) * (vc->dims[2] + 1) + index[1] * (vc->dims[2] + 1) +
                 index[2];
    double dp[3], ds[3];
    double dt = (exit_t - enter_t) / (double)vri->n_samples;
    double dvs[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3; i++) {
        dp[i] = (enter_t + 0.5 * dt) * v_dir[i] + v_pos[i];
        dp[i] -= (double)index[i] * vc->dds[i] + vc->left_edge[i];
        dp[i] *= vc->idds[i];
        ds[i] = v_dir[i] * vc->idds[i] * dt;
    }
    const int nch = vc->n_fields < 3 ? vc->n_fields : 3;
    for (int s = 0; s < vri->n_samples; s++) {
        for (int j = 0; j < vc->n_fields; j++)
            dvs[j] = orc_offset_interpolate(vc->dims, dp, vc->data[j] + offset);
        for (int j = 0; j < nch; j++) im->rgba[j] += dvs[j] * dt;
        for (int j = 0; j < 3; j++) dp[j] += ds[j];
    }
}

/* LightSourceRenderSampler.sample: image_samplers.pyx:548-590.  As the
 * volume renderer but without the mask test, with a gradient of field 0 per
 * sub-sample and FIT_eval_transfer_with_light. */
static void orc_light_source_sample(const orc_vc *vc, const double v_pos[3], const double v_dir[3],
                                    double enter_t, double exit_t, const int index[3], void *data)
{
    orc_acc *im = (orc_acc *)data;
    const orc_vra *vri = im->vra;
    if (!orc_in_range(vc, index)) return;
    int offset = index[0] * (vc->dims[1] + 1) * (vc->dims[2] + 1) + index[1] * (vc->dims[2] + 1) +
                 index[2];
    double dp[3], ds[3], grad[3];
    double dt = (exit_t - enter_t) / (double)vri->n_samples;
    double dvs[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3; i++) {
        dp[i] = (enter_t + 0.5 * dt) * v_dir[i] + v_pos[i];
        dp[i] -= (double)index[i] * vc->dds[i] + vc->left_edge[i];
        dp[i] *= vc->idds[i];
        ds[i] = v_dir[i] * vc->idds[i] * dt;
    }
    for (int s = 0; s < vri->n_samples; s++) {
        for (int j = 0; j < vc->n_fields; j++)
            dvs[j] = orc_offset_interpolate(vc->dims, dp, vc->data[j] + offset);
        orc_eval_gradient(vc->dims, dp, vc->data[0] + offset, grad);
        orc_fit_eval_transfer_with_light(dt, dvs, grad, vri->light_dir, vri->light_rgba, im->rgba,
                                         vri->n_fits, vri->fits, vri->field_table_ids,
                                         vri->grey_opacity);
        for (int j = 0; j < 3; j++) dp[j] += ds[j];
    }
}

/* Visit instrumentation: a 64-bit order-sensitive hash over
 * (brick id, i, j, k, bits(enter_t), bits(exit_t)) of every sampler call. */
static inline uint64_t orc_mix(uint64_t h, uint64_t w)
{
    h ^= w;
    h *= 0x100000001b3ULL;
    h ^= h >> 29;
    return h;
}
static inline uint64_t orc_dbits(double d)
{
    uint64_t u;
    memcpy(&u, &d, sizeof(u));
    return u;
}
static inline void orc_record(orc_acc *im, const int index[3], double enter_t, double exit_t)
{
    uint64_t h = im->hash;
    h = orc_mix(h, (uint64_t)(uint32_t)im->brick_id | ((uint64_t)(uint32_t)index[0] << 32));
    h = orc_mix(h, (uint64_t)(uint32_t)index[1] | ((uint64_t)(uint32_t)index[2] << 32));
    h = orc_mix(h, orc_dbits(enter_t));
    h = orc_mix(h, orc_dbits(exit_t));
    im->hash = h;
    im->visits += 1;
}
/* Samples of the measurement: one per projection visit with exit_t >
 * enter_t; n_samples per visit of an in-range cell for the samplers that
 * sub-sample, VolumeRenderSampler only where the mask is 1
 * (image_samplers.pyx:467-470). */
static void orc_count_subsamples(const orc_vc *vc, const int index[3], orc_acc *im, int masked)
{
    if (!orc_in_range(vc, index)) return;
    if (masked && vc->mask != NULL &&
        vc->mask[((size_t)index[0] * vc->dims[1] + index[1]) * vc->dims[2] + index[2]] != 1)
        return;
    im->samples += im->vra->n_samples;
}
static void orc_projection_sample_rec(const orc_vc *vc, const double v_pos[3], const double v_dir[3],
                                      double enter_t, double exit_t, const int index[3], void *data)
{
    orc_record((orc_acc *)data, index, enter_t, exit_t);
    if (exit_t > enter_t) ((orc_acc *)data)->samples += 1;
    orc_projection_sample(vc, v_pos, v_dir, enter_t, exit_t, index, data);
}
static void orc_interpolated_projection_sample_rec(const orc_vc *vc, const double v_pos[3],
                                                   const double v_dir[3], double enter_t,
                                                   double exit_t, const int index[3], void *data)
{
    orc_record((orc_acc *)data, index, enter_t, exit_t);
    orc_count_subsamples(vc, index, (orc_acc *)data, 0);
    orc_interpolated_projection_sample(vc, v_pos, v_dir, enter_t, exit_t, index, data);
}
static void orc_light_source_sample_rec(const orc_vc *vc, const double v_pos[3],
                                        const double v_dir[3], double enter_t, double exit_t,
                                        const int index[3], void *data)
{
    orc_record((orc_acc *)data, index, enter_t, exit_t);
    orc_count_subsamples(vc, index, (orc_acc *)data, 0);
    orc_light_source_sample(vc, v_pos, v_dir, enter_t, exit_t, index, data);
}
static void orc_volume_render_sample_rec(const orc_vc *vc, const double v_pos[3],
                                         const double v_dir[3], double enter_t, double exit_t,
                                         const int index[3], void *data)
{
    orc_record((orc_acc *)data, index, enter_t, exit_t);
    orc_count_subsamples(vc, index, (orc_acc *)data, 1);
    orc_volume_render_sample(vc, v_pos, v_dir, enter_t, exit_t, index, data);
}

/* ---------------------------------------------------------------------- */
/* Plane-parallel lens: lenses.pyx:22-55 and :179-192. */
typedef struct {
    double vp_pos[12]; /* inv_mat.ravel('F') (9) + back_center (3): lens.py:95-100 */
    double vp_dir[3];  /* box_vectors[2]: lens.py:104                             */
    double center[3];
    double bounds[4];
    double x_vec[3];
    double y_vec[3];
    double width[3];
    int nv[2];
    double pdx, pdy;
} orc_cam;

/* calculate_extent_plane_parallel: lenses.pyx:22-55.
 * Pinned by tests/test_oracle_lens.py (conservativeness vs the ray lattice
 * and rounding cases). */
static void orc_extent_plane_parallel(const orc_cam *cam, const orc_vc *vc, int64_t rv[4])
{
    double temp;
    const double *edges[2];
    double cx, cy;
    double extrema[4];
    edges[0] = vc->left_edge;
    edges[1] = vc->right_edge;
    extrema[0] = extrema[2] = 1e300;
    extrema[1] = extrema[3] = -1e300;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) {
                temp = edges[i][0] * cam->x_vec[0];
                temp += edges[j][1] * cam->x_vec[1];
                temp += edges[k][2] * cam->x_vec[2];
                if (temp < extrema[0]) extrema[0] = temp;
                if (temp > extrema[1]) extrema[1] = temp;
                temp = edges[i][0] * cam->y_vec[0];
                temp += edges[j][1] * cam->y_vec[1];
                temp += edges[k][2] * cam->y_vec[2];
                if (temp < extrema[2]) extrema[2] = temp;
                if (temp > extrema[3]) extrema[3] = temp;
            }
    cx = cy = 0.0;
    for (int i = 0; i < 3; i++) {
        cx += cam->center[i] * cam->x_vec[i];
        cy += cam->center[i] * cam->y_vec[i];
    }
    rv[0] = lrint((extrema[0] - cx - cam->bounds[0]) / cam->pdx);
    rv[1] = rv[0] + lrint((extrema[1] - extrema[0]) / cam->pdx);
    rv[2] = lrint((extrema[2] - cy - cam->bounds[2]) / cam->pdy);
    rv[3] = rv[2] + lrint((extrema[3] - extrema[2]) / cam->pdy);
}

/* generate_vector_info_plane_parallel: lenses.pyx:179-192. */
static inline void orc_vector_plane_parallel(const orc_cam *cam, int64_t vi, int64_t vj,
                                             const double width[3], double v_dir[3],
                                             double v_pos[3])
{
    double px, py;
    px = width[0] * (double)vi / ((double)cam->nv[0] - 1) - width[0] / 2.0;
    py = width[1] * (double)vj / ((double)cam->nv[1] - 1) - width[1] / 2.0;
    v_pos[0] = cam->vp_pos[0] * px + cam->vp_pos[3] * py + cam->vp_pos[9];
    v_pos[1] = cam->vp_pos[1] * px + cam->vp_pos[4] * py + cam->vp_pos[10];
    v_pos[2] = cam->vp_pos[2] * px + cam->vp_pos[5] * py + cam->vp_pos[11];
    for (int i = 0; i < 3; i++) v_dir[i] = cam->vp_dir[i];
}

/* ---------------------------------------------------------------------- */
/* Per-pixel lenses (perspective, stereo, fisheye, spherical): the sampler
 * reads ray origin and direction from (nv0, nv1, 3) arrays
 * (generate_vector_info_null, lenses.pyx:196-207).  The extent is the whole
 * image (calculate_extent_null, lenses.pyx:169-175) unless the lens is
 * perspective and camera data are given (image_samplers.pyx:98-99). */

/* calculate_extent_perspective: lenses.pyx:60-161.  camera_data rows:
 * position, width, east, north, normal (camera.py:280).  The reference
 * declares sight_vector_norm, sight_angle_cos, sight_length, dx and dy as C
 * float, so each is rounded to float32 where it is assigned; every other
 * quantity is double.  The vector helpers (vec3_ops.pxd) are the double
 * specialisations.  Pinned by tests/test_oracle_lens.py (every pixel whose
 * ray meets the brick lies in the clipped extent). */
static void orc_extent_perspective(const double *camera_data, int nv0, int nv1, const orc_vc *vc,
                                   int64_t rv[4])
{
    double cam_pos[3], cam_width[3], east[3], north[3], normal[3];
    double vertex[3], pos1[3], sight_vector[3], sight_center[3];
    double corners[3][8];
    float sight_vector_norm, sight_angle_cos, sight_length, dx, dy;
    int min_px = 32767, min_py = 32767, max_px = -32767, max_py = -32767;
    const double *le = vc->left_edge, *re = vc->right_edge;
    /* lenses.pyx:80-103: the corner order of the reference */
    const int cx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
    const int cy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    const int cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (int iv = 0; iv < 8; iv++) {
        corners[0][iv] = cx[iv] ? re[0] : le[0];
        corners[1][iv] = cy[iv] ? re[1] : le[1];
        corners[2][iv] = cz[iv] ? re[2] : le[2];
    }
    for (int i = 0; i < 3; i++) {
        cam_pos[i] = camera_data[0 * 3 + i];
        cam_width[i] = camera_data[1 * 3 + i];
        east[i] = camera_data[2 * 3 + i];
        north[i] = camera_data[3 * 3 + i];
        normal[i] = camera_data[4 * 3 + i];
    }
    for (int iv = 0; iv < 8; iv++) {
        vertex[0] = corners[0][iv];
        vertex[1] = corners[1][iv];
        vertex[2] = corners[2][iv];
        cam_width[1] = cam_width[0] * nv1 / nv0;
        for (int i = 0; i < 3; i++) sight_vector[i] = vertex[i] - cam_pos[i];
        for (int i = 0; i < 3; i++) sight_center[i] = cam_width[2] * normal[i] + cam_pos[i];
        sight_vector_norm = (float)sqrt(sight_vector[0] * sight_vector[0] +
                                        sight_vector[1] * sight_vector[1] +
                                        sight_vector[2] * sight_vector[2]);
        if (sight_vector_norm != 0)
            for (int i = 0; i < 3; i++) sight_vector[i] /= (double)sight_vector_norm;
        sight_angle_cos = (float)(sight_vector[0] * normal[0] + sight_vector[1] * normal[1] +
                                  sight_vector[2] * normal[2]);
        sight_angle_cos = (float)fmin(fmax((double)sight_angle_cos, -1.0), 1.0);
        if (acos((double)sight_angle_cos) < 0.5 * ORC_PI && sight_angle_cos != 0.0) {
            sight_length = (float)(cam_width[2] / (double)sight_angle_cos);
        } else {
            sight_length = (float)sqrt(cam_width[0] * cam_width[0] + cam_width[1] * cam_width[1]);
            const float c2 = sight_angle_cos * sight_angle_cos;
            sight_length = (float)((double)sight_length / sqrt(1.0 - (double)c2));
        }
        for (int i = 0; i < 3; i++) pos1[i] = (double)sight_length * sight_vector[i] + cam_pos[i];
        for (int i = 0; i < 3; i++) pos1[i] = pos1[i] - sight_center[i];
        dx = (float)(pos1[0] * east[0] + pos1[1] * east[1] + pos1[2] * east[2]);
        dy = (float)(pos1[0] * north[0] + pos1[1] * north[1] + pos1[2] * north[2]);
        const int px = (int)(nv0 * 0.5 + nv0 / cam_width[0] * (double)dx);
        const int py = (int)(nv1 * 0.5 + nv1 / cam_width[1] * (double)dy);
        if (px < min_px) min_px = px;
        if (px > max_px) max_px = px;
        if (py < min_py) min_py = py;
        if (py > max_py) max_py = py;
    }
    rv[0] = min_px > 0 ? min_px : 0;
    rv[1] = max_px < nv0 ? max_px : nv0;
    rv[2] = min_py > 0 ? min_py : 0;
    rv[3] = max_py < nv1 ? max_py : nv1;
}

/* ---------------------------------------------------------------------- */
/* Public oracle entry points (ctypes). */

typedef struct {
    int32_t n_fits;
    int32_t nbins[6];
    double bounds[12];       /* per table: (b0, b1)                     */
    const double *y[6];      /* per table: nbins values                 */
    int32_t field_ids[6];
    int32_t weight_field_ids[6];
    int32_t weight_table_ids[6];
    int32_t field_table_ids[6];
    int32_t grey_opacity;
    double light_dir[3];  /* LightSourceRenderSampler only */
    double light_rgba[4];
} orc_tf_desc;

/*
 * orc_cast_lens: ImageSampler.__call__ -> cast_through_kdtree
 * (image_samplers.pyx:132-197) for each brick of `order` in sequence.
 * kind 0 = ProjectionSampler, 1 = VolumeRenderSampler,
 * 2 = InterpolatedProjectionSampler, 3 = LightSourceRenderSampler.
 * lens 0: plane-parallel (vp_pos/vp_dir/center/bounds/x_vec/y_vec/width);
 * lens 1: per-pixel rays pix_pos/pix_dir (nv0, nv1, 3), with the perspective
 * extent when camera_data (5 x 3) is given, else the whole image.  OpenMP: static schedule, chunk 100, as the
 * reference's prange.  kind 0 = ProjectionSampler, 1 = VolumeRenderSampler.
 * image: float64 (nv0, nv1, 4), read-modify-write.  zbuffer: (nv0, nv1) or NULL
 * (NULL means ones, utils.py:136).  visits/hashes: per-pixel instrumentation
 * arrays or NULL; samples: per-pixel sample counts (orc_count_subsamples) or
 * NULL, filled only with visits/hashes.  Returns 0, or -1 on bad arguments.
 */
/* Rows of the image a cast covers (default: all).  Only for timing the
 * oracle on a bounded sample of a large image: the pixels of rows outside
 * [orc_row0, orc_row1) are left as they are. */
static int orc_row0 = 0, orc_row1 = 0x7fffffff;

void orc_set_rows(int r0, int r1)
{
    orc_row0 = r0 < 0 ? 0 : r0;
    orc_row1 = r1;
}

int orc_cast_lens(int kind, int lens, const double *pix_pos, const double *pix_dir,
                  const double *camera_data, const double *vp_pos, const double *vp_dir,
                  const double *center, const double *bounds, const double *x_vec,
                  const double *y_vec, const double *width, int nv0, int nv1, int n_bricks,
                  const double *les, const double *res, const int32_t *dims, int n_fields,
                  const double *const *field_ptrs, const uint8_t *const *mask_ptrs, int n_order,
                  const int32_t *order, const orc_tf_desc *tf, int n_samples, double *image,
                  const double *zbuffer, int num_threads, int64_t *visits, uint64_t *hashes,
                  int64_t *samples)
{
    if (nv0 <= 0 || nv1 <= 0 || n_fields <= 0 || n_fields > 6) return -1;
    if (lens != 0 && (pix_pos == NULL || pix_dir == NULL)) return -1;
    if ((kind == 1 || kind == 3) && (tf == NULL || tf->n_fits > 6 || n_samples <= 0)) return -1;
    orc_cam cam;
    memcpy(cam.vp_pos, vp_pos, sizeof(cam.vp_pos));
    memcpy(cam.vp_dir, vp_dir, sizeof(cam.vp_dir));
    memcpy(cam.center, center, sizeof(cam.center));
    memcpy(cam.bounds, bounds, sizeof(cam.bounds));
    memcpy(cam.x_vec, x_vec, sizeof(cam.x_vec));
    memcpy(cam.y_vec, y_vec, sizeof(cam.y_vec));
    memcpy(cam.width, width, sizeof(cam.width));
    cam.nv[0] = nv0;
    cam.nv[1] = nv1;
    /* image_samplers.pyx:127-128 */
    cam.pdx = (cam.bounds[1] - cam.bounds[0]) / (double)cam.nv[0];
    cam.pdy = (cam.bounds[3] - cam.bounds[2]) / (double)cam.nv[1];

    orc_fit fits[6];
    orc_vra vra;
    memset(&vra, 0, sizeof(vra));
    if (kind == 1 || kind == 3) {
        for (int i = 0; i < 3; i++) vra.light_dir[i] = tf->light_dir[i];
        for (int i = 0; i < 4; i++) vra.light_rgba[i] = tf->light_rgba[i];
        vra.n_fits = tf->n_fits;
        vra.n_samples = n_samples;
        vra.grey_opacity = tf->grey_opacity;
        for (int i = 0; i < tf->n_fits; i++)
            orc_fit_init(&fits[i], tf->nbins[i], tf->y[i], tf->bounds[2 * i], tf->bounds[2 * i + 1],
                         tf->field_ids[i], tf->weight_field_ids[i], tf->weight_table_ids[i]);
        for (int i = 0; i < 6; i++) vra.field_table_ids[i] = tf->field_table_ids[i];
        vra.fits = fits;
    }
    const int rec = (visits != NULL && hashes != NULL);
    orc_sampler fn;
    if (kind == 0)
        fn = rec ? orc_projection_sample_rec : orc_projection_sample;
    else if (kind == 2)
        fn = rec ? orc_interpolated_projection_sample_rec : orc_interpolated_projection_sample;
    else if (kind == 3)
        fn = rec ? orc_light_source_sample_rec : orc_light_source_sample;
    else
        fn = rec ? orc_volume_render_sample_rec : orc_volume_render_sample;
    if (kind == 2) {
        if (n_samples <= 0) return -1;
        vra.n_samples = n_samples;
    }

    if (num_threads <= 0) num_threads = omp_get_max_threads();

    for (int ob = 0; ob < n_order; ob++) {
        int b = order[ob];
        if (b < 0 || b >= n_bricks) return -1;
        orc_vc vc;
        orc_vc_init(&vc, les + 3 * b, res + 3 * b, dims + 3 * b);
        vc.n_fields = n_fields;
        vc.data = field_ptrs + (size_t)b * n_fields;
        vc.mask = mask_ptrs ? mask_ptrs[b] : NULL;
        int64_t it[4];
        if (lens == 0) {
            orc_extent_plane_parallel(&cam, &vc, it);
        } else if (camera_data != NULL) {
            orc_extent_perspective(camera_data, nv0, nv1, &vc, it);
        } else {
            it[0] = 0;
            it[1] = nv0;
            it[2] = 0;
            it[3] = nv1;
        }
        /* image_samplers.pyx:160-166 */
        it[0] = y_iclip(it[0] - 1, 0, cam.nv[0]);
        it[1] = y_iclip(it[1] + 1, 0, cam.nv[0]);
        it[2] = y_iclip(it[2] - 1, 0, cam.nv[1]);
        it[3] = y_iclip(it[3] + 1, 0, cam.nv[1]);
        if (it[0] < orc_row0) it[0] = orc_row0;
        if (it[1] > orc_row1) it[1] = orc_row1;
        int64_t nx = it[1] - it[0];
        int64_t ny = it[3] - it[2];
        int64_t size = nx * ny;
        if (size <= 0) continue;
        double wid[3] = {cam.width[0], cam.width[1], cam.width[2]};
#pragma omp parallel for schedule(static, 100) num_threads(num_threads)
        for (int64_t j = 0; j < size; j++) {
            double v_pos[3], v_dir[3];
            orc_acc acc;
            int64_t vj = j % ny;
            int64_t vi = (j - vj) / ny + it[0];
            vj = vj + it[2];
            size_t pix = (size_t)vi * (size_t)nv1 + (size_t)vj;
            if (lens == 0) {
                orc_vector_plane_parallel(&cam, vi, vj, wid, v_dir, v_pos);
            } else {
                /* generate_vector_info_null: lenses.pyx:196-207 */
                for (int i = 0; i < 3; i++) {
                    v_pos[i] = pix_pos[pix * 3 + i];
                    v_dir[i] = pix_dir[pix * 3 + i];
                }
            }
            for (int c = 0; c < 4; c++) acc.rgba[c] = image[pix * 4 + c];
            acc.vra = &vra;
            acc.brick_id = b;
            acc.visits = rec ? visits[pix] : 0;
            acc.hash = rec ? hashes[pix] : 0;
            acc.samples = (rec && samples) ? samples[pix] : 0;
            double max_t = y_fclip(zbuffer ? zbuffer[pix] : 1.0, 0.0, 1.0);
            orc_walk_volume(&vc, v_pos, v_dir, fn, &acc, max_t);
            for (int c = 0; c < 4; c++) image[pix * 4 + c] = acc.rgba[c];
            if (rec) {
                visits[pix] = acc.visits;
                hashes[pix] = acc.hash;
                if (samples) samples[pix] = acc.samples;
            }
        }
    }
    if (kind == 1 || kind == 3)
        for (int i = 0; i < tf->n_fits; i++) {
            free(fits[i].d0);
            free(fits[i].dy);
        }
    return 0;
}

int orc_cast(int kind, const double *vp_pos, const double *vp_dir, const double *center,
             const double *bounds, const double *x_vec, const double *y_vec, const double *width,
             int nv0, int nv1, int n_bricks, const double *les, const double *res,
             const int32_t *dims, int n_fields, const double *const *field_ptrs,
             const uint8_t *const *mask_ptrs, int n_order, const int32_t *order,
             const orc_tf_desc *tf, int n_samples, double *image, const double *zbuffer,
             int num_threads, int64_t *visits, uint64_t *hashes, int64_t *samples)
{
    return orc_cast_lens(kind, 0, NULL, NULL, NULL, vp_pos, vp_dir, center, bounds, x_vec, y_vec,
                         width, nv0, nv1, n_bricks, les, res, dims, n_fields, field_ptrs, mask_ptrs,
                         n_order, order, tf, n_samples, image, zbuffer, num_threads, visits, hashes,
                         samples);
}

/* Clipped perspective extents (lenses.pyx:60-161 + image_samplers.pyx:159-166).
 * out: n_bricks x 4 int64. */
int orc_extents_perspective(const double *camera_data, int nv0, int nv1, int n_bricks,
                            const double *les, const double *res, const int32_t *dims, int64_t *out)
{
    for (int b = 0; b < n_bricks; b++) {
        orc_vc vc;
        orc_vc_init(&vc, les + 3 * b, res + 3 * b, dims + 3 * b);
        int64_t *it = out + 4 * b;
        orc_extent_perspective(camera_data, nv0, nv1, &vc, it);
        it[0] = y_iclip(it[0] - 1, 0, nv0);
        it[1] = y_iclip(it[1] + 1, 0, nv0);
        it[2] = y_iclip(it[2] - 1, 0, nv1);
        it[3] = y_iclip(it[3] + 1, 0, nv1);
    }
    return 0;
}

/* Per-brick pixel extents after the +-1 clip (image_samplers.pyx:159-166),
 * for checking the product's host-side culling. out: n_bricks x 4 int64. */
int orc_extents(const double *center, const double *bounds, const double *x_vec,
                const double *y_vec, int nv0, int nv1, int n_bricks, const double *les,
                const double *res, const int32_t *dims, int64_t *out)
{
    orc_cam cam;
    memset(&cam, 0, sizeof(cam));
    memcpy(cam.center, center, sizeof(cam.center));
    memcpy(cam.bounds, bounds, sizeof(cam.bounds));
    memcpy(cam.x_vec, x_vec, sizeof(cam.x_vec));
    memcpy(cam.y_vec, y_vec, sizeof(cam.y_vec));
    cam.nv[0] = nv0;
    cam.nv[1] = nv1;
    cam.pdx = (cam.bounds[1] - cam.bounds[0]) / (double)cam.nv[0];
    cam.pdy = (cam.bounds[3] - cam.bounds[2]) / (double)cam.nv[1];
    for (int b = 0; b < n_bricks; b++) {
        orc_vc vc;
        orc_vc_init(&vc, les + 3 * b, res + 3 * b, dims + 3 * b);
        int64_t *it = out + 4 * b;
        orc_extent_plane_parallel(&cam, &vc, it);
        it[0] = y_iclip(it[0] - 1, 0, cam.nv[0]);
        it[1] = y_iclip(it[1] + 1, 0, cam.nv[0]);
        it[2] = y_iclip(it[2] - 1, 0, cam.nv[1]);
        it[3] = y_iclip(it[3] + 1, 0, cam.nv[1]);
    }
    return 0;
}

/* Visit log of one ray through one brick: walk_volume with a sampler that
 * records (i, j, k) and (enter_t, exit_t).  Returns walk_volume's hit count;
 * at most `cap` visits are written. */
typedef struct {
    int32_t *idx;
    double *t;
    int cap;
    int n;
} orc_log;

static void orc_log_sample(const orc_vc *vc, const double v_pos[3], const double v_dir[3],
                           double enter_t, double exit_t, const int index[3], void *data)
{
    (void)vc;
    (void)v_pos;
    (void)v_dir;
    orc_log *lg = (orc_log *)data;
    if (lg->n < lg->cap) {
        lg->idx[3 * lg->n + 0] = index[0];
        lg->idx[3 * lg->n + 1] = index[1];
        lg->idx[3 * lg->n + 2] = index[2];
        lg->t[2 * lg->n + 0] = enter_t;
        lg->t[2 * lg->n + 1] = exit_t;
    }
    lg->n += 1;
}

int orc_walk_log(const double *le, const double *re, const int32_t *dims, const double *v_pos,
                 const double *v_dir, double max_t, int32_t *idx_out, double *t_out, int cap)
{
    orc_vc vc;
    orc_vc_init(&vc, le, re, dims);
    vc.n_fields = 0;
    vc.data = NULL;
    vc.mask = NULL;
    orc_log lg = {idx_out, t_out, cap, 0};
    return orc_walk_volume(&vc, v_pos, v_dir, orc_log_sample, &lg, max_t);
}

/* FIT_get_value for a vector of inputs (single table, no weights). */
int orc_fit_values(int nbins, const double *y, double b0, double b1, int n, const double *v,
                   double *out)
{
    orc_fit fit;
    orc_fit_init(&fit, nbins, y, b0, b1, 0, -1, -1);
    for (int i = 0; i < n; i++) {
        double dvs[6] = {v[i], 0, 0, 0, 0, 0};
        out[i] = orc_fit_get_value(&fit, dvs);
    }
    free(fit.d0);
    free(fit.dy);
    return 0;
}

/* offset_interpolate on a (d0+1, d1+1, d2+1) vertex array for n points
 * given as (cell index, dp) pairs. */
int orc_trilinear(const int32_t *dims, const double *vdata, int n, const int32_t *cells,
                  const double *dps, double *out)
{
    int ds[3] = {dims[0], dims[1], dims[2]};
    for (int p = 0; p < n; p++) {
        const int32_t *c = cells + 3 * p;
        int offset = c[0] * (ds[1] + 1) * (ds[2] + 1) + c[1] * (ds[2] + 1) + c[2];
        out[p] = orc_offset_interpolate(ds, dps + 3 * p, vdata + offset);
    }
    return 0;
}

/* eval_gradient at n positions of cells of one vertex-centred brick. */
int orc_gradients(const int32_t *dims, const double *vdata, int n, const int32_t *cells,
                  const double *dps, double *out)
{
    int ds[3] = {dims[0], dims[1], dims[2]};
    for (int p = 0; p < n; p++) {
        const int32_t *c = cells + 3 * p;
        int offset = c[0] * (ds[1] + 1) * (ds[2] + 1) + c[1] * (ds[2] + 1) + c[2];
        orc_eval_gradient(ds, dps + 3 * p, vdata + offset, out + 3 * p);
    }
    return 0;
}

int orc_max_threads(void) { return omp_get_max_threads(); }
