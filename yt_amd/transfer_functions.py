"""Transfer-function objects (host side), restating
yt/visualization/volume_rendering/transfer_functions.py:
  TransferFunction                 :34-170 (add_gaussian, add_line, add_step)
  MultiVariateTransferFunction     :249-340
  ColorTransferFunction            :363-445
  ProjectionTransferFunction       :921-961
They carry exactly the attributes VolumeRenderSampler reads
(image_samplers.pyx:429-446): n_field_tables, tables[i].y,
tables[i].x_bounds, field_ids, weight_field_ids, weight_table_ids,
field_table_ids, grey_opacity.
"""

import numpy as np


class TransferFunction:
    def __init__(self, x_bounds, nbins=256):
        self.pass_through = 0
        self.nbins = nbins
        x_bounds = [np.float64(xb) for xb in x_bounds]
        self.x_bounds = x_bounds
        self.x = np.linspace(x_bounds[0], x_bounds[1], nbins, dtype="float64")
        self.y = np.zeros(nbins, dtype="float64")
        self.grad_field = -1
        self.light_source_v = self.light_source_c = np.zeros(3, "float64")
        self.features = []

    def add_gaussian(self, location, width, height):
        """f(x) = h exp(-(x-x0)^2 / w), max-combined, clipped at 0 (:46-80)."""
        vals = height * np.exp(-((self.x - location) ** 2.0) / width)
        self.y = np.clip(np.maximum(vals, self.y), 0.0, np.inf)
        self.features.append(("gaussian", location, width, height))

    def add_line(self, start, stop):
        x0, y0 = start
        x1, y1 = stop
        slope = (y1 - y0) / (x1 - x0)
        vals = slope * (self.x - x0) + y0
        vals[~((self.x >= x0) & (self.x <= x1))] = 0.0
        self.y = np.clip(np.maximum(vals, self.y), 0.0, np.inf)
        self.features.append(("line", start, stop))

    def add_step(self, start, stop, value):
        vals = np.zeros(self.x.shape, "float64")
        vals[(self.x >= start) & (self.x <= stop)] = value
        self.y = np.clip(np.maximum(vals, self.y), 0.0, np.inf)
        self.features.append(("step", start, stop, value))


class MultiVariateTransferFunction:
    def __init__(self, grey_opacity=False):
        self.n_field_tables = 0
        self.tables = []
        self.field_ids = [0] * 6
        self.weight_field_ids = [-1] * 6
        self.field_table_ids = [0] * 6
        self.weight_table_ids = [-1] * 6
        self.grad_field = -1
        self.light_source_v = self.light_source_c = np.zeros(3, "float64")
        self.grey_opacity = grey_opacity

    def add_field_table(self, table, field_id, weight_field_id=-1, weight_table_id=-1):
        self.tables.append(table)
        self.field_ids[self.n_field_tables] = field_id
        self.weight_field_ids[self.n_field_tables] = weight_field_id
        self.weight_table_ids[self.n_field_tables] = weight_table_id
        self.n_field_tables += 1

    def link_channels(self, table_id, channels=0):
        if np.ndim(channels) == 0:
            channels = [channels]
        for c in channels:
            self.field_table_ids[c] = table_id


class ColorTransferFunction(MultiVariateTransferFunction):
    """Four tables on field 0; R, G, B weighted by the alpha table (:385-408)."""

    def __init__(self, x_bounds, nbins=256, grey_opacity=False):
        MultiVariateTransferFunction.__init__(self)
        x_bounds = [np.float64(xb) for xb in x_bounds]
        self.x_bounds = x_bounds
        self.nbins = nbins
        self.red = TransferFunction(x_bounds, nbins)
        self.green = TransferFunction(x_bounds, nbins)
        self.blue = TransferFunction(x_bounds, nbins)
        self.alpha = TransferFunction(x_bounds, nbins)
        self.funcs = (self.red, self.green, self.blue, self.alpha)
        self.grey_opacity = grey_opacity
        self.features = []
        for i, tf in enumerate(self.funcs[:3]):
            self.add_field_table(tf, 0, weight_table_id=3)
            self.link_channels(i, i)
        self.add_field_table(self.funcs[3], 0)
        self.link_channels(3, 3)

    def add_gaussian(self, location, width, height):
        if len(height) != 4:
            raise ValueError("height must have 4 entries (r, g, b, alpha)")
        for tf, v in zip(self.funcs, height):
            tf.add_gaussian(location, width, v)
        self.features.append(("gaussian", location, width, tuple(height)))

    def add_step(self, start, stop, value):
        for tf, v in zip(self.funcs, value):
            tf.add_step(start, stop, v)


class ProjectionTransferFunction(MultiVariateTransferFunction):
    """Selects the projection sampler (render_source.py:295-296)."""

    def __init__(self, x_bounds=(-1e60, 1e60), n_fields=1):
        if n_fields > 3:
            raise NotImplementedError(f"supplied ${n_fields} but n_fields > 3 not implemented.")
        MultiVariateTransferFunction.__init__(self)
        x_bounds = [np.float64(xb) for xb in x_bounds]
        self.x_bounds = x_bounds
        self.nbins = 2
        self.linear_mapping = TransferFunction(x_bounds, 2)
        self.linear_mapping.pass_through = 1
        self.link_channels(0, [0, 1, 2])
        for i in range(n_fields):
            self.add_field_table(self.linear_mapping, i)
            self.link_channels(i, i)
        self.link_channels(n_fields, [3, 4, 5])


def gaussian_ctf(bounds, n_gauss=8, nbins=256, alpha_range=(-3.0, 0.0), colors=None):
    """The 8-Gaussian ColorTransferFunction of the AMR and batched workloads:
    evenly spaced centres over `bounds`, w = 0.01 * range^2, RGB heights from
    a fixed table, alpha heights logspace(alpha_range, n_gauss)."""
    lo, hi = float(bounds[0]), float(bounds[1])
    tf = ColorTransferFunction((lo, hi), nbins=nbins)
    rng = hi - lo
    if colors is None:
        base = np.array([
            [0.9, 0.1, 0.1], [0.9, 0.5, 0.1], [0.9, 0.9, 0.1], [0.3, 0.9, 0.2],
            [0.1, 0.8, 0.8], [0.1, 0.4, 0.9], [0.5, 0.2, 0.9], [0.9, 0.2, 0.7],
        ])
        colors = base[np.arange(n_gauss) % len(base)]
    alphas = np.logspace(alpha_range[0], alpha_range[1], n_gauss)
    locs = lo + rng * (np.arange(n_gauss) + 0.5) / n_gauss
    for i in range(n_gauss):
        r, g, b = colors[i]
        tf.add_gaussian(locs[i], 0.01 * rng * rng, [r, g, b, alphas[i]])
    return tf
