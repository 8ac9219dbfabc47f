"""Shared builders for tests: seeded datasets, cameras and oracle inputs."""

import numpy as np

from yt_amd import data, render
from yt_amd.transfer_functions import gaussian_ctf


def oracle_proj_bricks(ds, fields, weight=None):
    """Cell-centred, child-masked blocks in index order, as
    off_axis_projection hands them to ProjectionSampler."""
    out = []
    for g, m in ds.blocks():
        mm = m.astype("float64")
        if weight is None:
            d = [np.asarray(g[f], dtype="float64") * mm for f in fields]
        else:
            w = np.asarray(g[weight], dtype="float64")
            d = [np.asarray(g[fields[0]], dtype="float64") * w * mm, w * mm]
        out.append(dict(le=g.LeftEdge, re=g.RightEdge, dims=g.ActiveDimensions, data=d,
                        mask=None))
    return out


def oracle_vr_bricks(vol):
    """The kd bricks of a ResidentVolume, in the same (depth-first) index
    order as the device dataset, with the float64 brick data (the FP64
    reference the float32 device path is measured against)."""
    out = []
    for node in vol.leaves:
        pg = vol.kd.get_brick_data(node)
        out.append(dict(le=pg.LeftEdge, re=pg.RightEdge, dims=pg.dims,
                        data=[np.asarray(d, dtype=np.float64) for d in pg.my_data],
                        mask=pg.source_mask))
        node.data = None
    return out


def tf_dict(tf):
    n = tf.n_field_tables
    return dict(tables=[(tf.tables[i].y, tf.tables[i].x_bounds[0], tf.tables[i].x_bounds[1])
                        for i in range(n)],
                field_ids=tf.field_ids, weight_field_ids=tf.weight_field_ids,
                weight_table_ids=tf.weight_table_ids, field_table_ids=tf.field_table_ids,
                grey_opacity=int(tf.grey_opacity))


def small_amr(root_n=32, levels=2, seed=7):
    gd = data.make_amr_hierarchy(root_n=root_n, levels=levels, seed_density=seed,
                                 seed_flag=seed + 1, percentile=97.0, block=8, max_block=16)
    return data.load_amr_grids(gd, [root_n] * 3)


def vr_tf_for(ds, field="density", nbins=256):
    lo = min(np.log10(g[field].min()) for g in ds.grids)
    hi = max(np.log10(g[field].max()) for g in ds.grids)
    return gaussian_ctf((lo, hi), n_gauss=8, nbins=nbins)


def uniform_lognormal(n, seed=1001, nprocs=1, psize=None, with_temperature=False):
    rho = data.lognormal_density(n, seed=seed)
    fields = {"density": rho}
    if with_temperature:
        fields["temperature"] = data.temperature_from_density(rho, seed=seed + 1)
    return data.load_uniform_grid(fields, nprocs=nprocs, psize=psize)


C0_NORMAL = (0.3, 0.5, 0.81)


def c0_params(res=256, normal=C0_NORMAL, width=1.0, center=(0.5, 0.5, 0.5), north=None):
    return render.camera_params(center, normal, width, res, north)
