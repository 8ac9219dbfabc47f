"""Host-side composition and image output (no GPU): ZBuffer merging
(zbuffer_array.py), ImageArray rescale / background / sigma clip
(image_array.py) and the PNG writer (image_writer.py write_bitmap)."""

import struct
import zlib

import numpy as np

from yt_amd.scene import ZBuffer, add_background_color, clipping_value, rescale, write_bitmap


def test_zbuffer_keeps_nearer():
    """zbuffer_array.py's docstring example and a mixed-depth case."""
    shape = (16, 16)
    rng = np.random.default_rng(3)
    b1 = ZBuffer(rng.random(shape), np.ones(shape))
    b2 = ZBuffer(rng.random(shape), np.zeros(shape))
    c = b1 + b2
    assert np.all(c.rgba == b2.rgba) and np.all(c.z == b2.z) and c == b2
    z1 = rng.random(shape)
    z2 = rng.random(shape)
    r1, r2 = rng.random(shape + (4,)), rng.random(shape + (4,))
    d = ZBuffer(r1, z1) + ZBuffer(r2, z2)
    near1 = z1 < z2
    assert np.array_equal(d.rgba[near1], r1[near1]) and np.array_equal(d.rgba[~near1], r2[~near1])
    assert np.array_equal(d.z, np.minimum(z1, z2))


def test_rescale_background_clip():
    im = np.zeros((8, 10, 4))
    for k in range(4):
        im[:, :, k] = np.linspace(0.0, 0.3 * (k + 1), 10)[None, :]
    r = rescale(im)
    assert np.isclose(r[:, :, :3].sum(axis=2).max(), 1.0 * r[:, :, :3].sum(axis=2).max())
    assert r[:, :, 3].max() == 1.0 and r.min() >= 0.0
    assert np.allclose(r[:, :, :3], np.clip(im[:, :, :3] / im[:, :, :3].sum(axis=2).max(), 0, 1))
    b = add_background_color(r, "black")
    assert np.allclose(b[:, :, 3], 1.0)
    assert np.allclose(b[:, :, 0], r[:, :, 0] * r[:, :, 3])
    nz = r[:, :, :3][r[:, :, :3].nonzero()]
    assert np.isclose(clipping_value(r, 2.0), nz.mean() + 2.0 * nz.std())


def _decode_png(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        tag = data[pos + 4:pos + 8]
        payload = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(tag + payload) & 0xFFFFFFFF
        chunks[tag] = chunks.get(tag, b"") + payload
        pos += 12 + n
    w, h, depth, ctype = struct.unpack(">IIBB", chunks[b"IHDR"][:10])
    raw = zlib.decompress(chunks[b"IDAT"])
    rows = [raw[y * (4 * w + 1) + 1:(y + 1) * (4 * w + 1)] for y in range(h)]
    return np.frombuffer(b"".join(rows), np.uint8).reshape(h, w, 4), depth, ctype


def test_png_round_trip():
    rng = np.random.default_rng(5)
    im = rng.random((12, 7, 4))
    arr = write_bitmap(im, None)
    dec, depth, ctype = _decode_png(arr)
    assert depth == 8 and ctype == 6
    want = np.concatenate([(np.clip(im[:, :, :3] / im[:, :, :3].max(), 0, 1) * 255).astype("uint8"),
                           (255 * im[:, :, 3]).astype("uint8")[..., None]], axis=-1)
    assert np.array_equal(dec, want)


def test_volume_source_transfer_function_rules():
    """render_source.py:104-134, 280-302: a ProjectionTransferFunction makes
    the source a linear "projection" source; other types are refused; a
    weight field is the second kd field with the same log flag."""
    import pytest

    from yt_amd.scene import VolumeSource
    from yt_amd.transfer_functions import (
        ColorTransferFunction,
        MultiVariateTransferFunction,
        ProjectionTransferFunction,
    )

    s = VolumeSource(None, "density", transfer_function=ColorTransferFunction((0.0, 1.0)))
    assert s.sampler_type == "volume-render" and s.log_field is True
    assert s._fields() == (["density"], [True])
    s.set_weight_field("cell_mass")
    assert s._fields() == (["density", "cell_mass"], [True, True])
    s.transfer_function = ProjectionTransferFunction(n_fields=2)
    assert s.sampler_type == "projection" and s.log_field is False
    assert s._fields() == (["density", "cell_mass"], [False, False])
    with pytest.raises(RuntimeError):
        s.transfer_function = MultiVariateTransferFunction()
    with pytest.raises(NotImplementedError):
        ProjectionTransferFunction(n_fields=4)


def test_zbuffer_ties_and_columns():
    """Equal depths clear the pixel in 2-d depth maps; one-column depth maps
    blend with weights 1 / 0, the second buffer winning ties
    (zbuffer_array.py:44-57)."""
    z = np.array([[0.2, 0.5], [0.7, 0.1]])
    a = ZBuffer(np.full((2, 2, 4), 1.0), z.copy())
    b = ZBuffer(np.full((2, 2, 4), 2.0), np.array([[0.3, 0.5], [0.6, 0.1]]))
    c = a + b
    assert np.array_equal(c.rgba[:, :, 0], [[1.0, 0.0], [2.0, 0.0]])
    zc = np.array([[0.2], [0.5], [0.9]])
    d = ZBuffer(np.full((3, 1, 4), 1.0), zc) + ZBuffer(np.full((3, 1, 4), 2.0),
                                                      np.array([[0.4], [0.5], [0.1]]))
    assert np.array_equal(d.rgba[:, 0, 0], [1.0, 2.0, 2.0])
    a += b
    assert a == c
    a.paint((0, 0), np.full(4, 7.0), 0.05)
    a.paint((1, 1), np.full(4, 9.0), 0.5)
    assert a.rgba[0, 0, 0] == 7.0 and a.z[0, 0] == 0.05 and a.rgba[1, 1, 0] == 0.0
