"""HDRI environment: the Radiance RGBE loader and 32 x 16 LUT hydration of
the render entry points (metal_bridge.mm:2163-2300, host code in the HIP
library, no GPU work) and the LUT sampler of environment() in the oracle
(Shaders.metal:5796-5844).

Pins:
  - decoding follows the Radiance format definition: a pixel (r, g, b, e)
    is (c + 0.5) 2^(e - 136), e = 0 is black; run and literal packets; the
    Y and X signs of the resolution line orient the image;
  - each LUT texel is the mean of its integer source box, rounded to FP16:
    numpy's float16 (round to nearest even) wherever the float32 mean is not
    halfway between two halves, away from zero at the halfway points;
  - sampling: a direction at a texel centre returns that texel; the u = 0
    seam and the midpoint between two rows average their neighbours; a yaw
    of one column (world[2] = 360/32 degrees) shifts the lookup by one column.
"""
import ctypes as C
import math

import numpy as np
import pytest

from fractalpathtracer_amd import api, ffi, scenes
from oracle import bind


def _rle_channel(vals):
    """Adaptive RLE of one channel of a scanline: runs of 3+ equal bytes as
    (128 + n, v), the rest as literal packets of up to 128 bytes."""
    out, i, n = bytearray(), 0, len(vals)
    while i < n:
        j = i
        while j < n and j - i < 127 and vals[j] == vals[i]:
            j += 1
        if j - i >= 3:
            out += bytes([128 + (j - i), vals[i]])
            i = j
            continue
        k = i
        while k < n and k - i < 128:
            r = k
            while r < n and r - k < 3 and vals[r] == vals[k]:
                r += 1
            if r - k >= 3:
                break
            k += 1
        out += bytes([k - i]) + bytes(vals[i:k])
        i = k
    return bytes(out)


def write_hdr(path, rgbe, ysign="-", xsign="+", fmt=True):
    """Radiance file from an (h, w, 4) uint8 RGBE array stored in memory
    top-down, left-right; the file rows/columns follow ysign/xsign."""
    h, w, _ = rgbe.shape
    body = bytearray(b"#?RADIANCE\n" + (b"FORMAT=32-bit_rle_rgbe\n" if fmt else b"") + b"\n")
    body += f"{ysign}Y {h} {xsign}X {w}\n".encode()
    rows = rgbe if ysign == "-" else rgbe[::-1]
    for row in rows:
        row = row if xsign == "+" else row[::-1]
        body += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            body += _rle_channel(list(row[:, c]))
    path.write_bytes(bytes(body))
    return str(path)


def decode(rgbe):
    """Radiance definition: (c + 0.5) 2^(e - 136), black where e = 0."""
    e = rgbe[..., 3].astype(np.int32)
    scale = np.where(e == 0, 0.0, np.ldexp(1.0, e - 136)).astype(np.float32)
    return ((rgbe[..., :3].astype(np.float32) + np.float32(0.5)) * scale[..., None]).astype(np.float32)


def half_round_away(v):
    """float32 -> FP16 bits: nearest, ties away from zero (normal range)."""
    v = np.float32(v)
    h = np.float16(v)
    lo = np.float16(v)
    # halfway: the float32 value is exactly between two adjacent halves
    nxt = np.nextafter(h, np.float16(np.inf) if v >= 0 else np.float16(-np.inf))
    prv = np.nextafter(h, np.float16(-np.inf) if v >= 0 else np.float16(np.inf))
    for a, b in ((h, nxt), (prv, h)):
        if np.float32(a) != np.float32(b) and (np.float64(a) + np.float64(b)) / 2 == np.float64(v):
            lo = a if abs(np.float64(a)) > abs(np.float64(b)) else b
    return int(np.array(lo, np.float16).view(np.uint16))


def box_lut(img):
    """Reference box downsampling: texel (x, y) averages source pixels
    [x w / 32, max(x0 + 1, (x + 1) w / 32)) x the same in y, summed in row
    order in float32."""
    h, w, _ = img.shape
    out = np.zeros((16, 32, 3), np.uint16)
    for ty in range(16):
        y0 = ty * h // 16
        y1 = min(max(y0 + 1, (ty + 1) * h // 16), h)
        for tx in range(32):
            x0 = tx * w // 32
            x1 = min(max(x0 + 1, (tx + 1) * w // 32), w)
            acc = np.zeros(3, np.float32)
            n = 0
            for sy in range(y0, y1):
                for sx in range(x0, x1):
                    acc = (acc + img[sy, sx]).astype(np.float32)
                    n += 1
            mean = (acc / np.float32(n)).astype(np.float32)
            out[ty, tx] = [half_round_away(m) for m in mean]
    return out


def _hdri_config(path, mode=2.0):
    c = scenes.default_config()
    scenes.set_hdri_path(c, path)
    c.world[0] = mode
    return c


def _lut(cfg):
    return np.frombuffer(bytes(cfg.hdri_lut), np.uint16).reshape(16, 32, 3)


def _random_rgbe(rng, h, w):
    rgbe = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    rgbe[..., 3] = rng.integers(124, 143, (h, w), dtype=np.uint8)  # normal FP16 range
    rgbe[rng.random((h, w)) < 0.05, 3] = 0  # some black pixels
    rgbe[2, :, :] = rgbe[2, 0, :]  # a run-length row
    return rgbe


@pytest.mark.parametrize("shape", [(32, 64), (20, 40), (4, 8), (17, 45)])
def test_lut_is_the_box_mean_of_the_decoded_image(tmp_path, shape):
    rng = np.random.default_rng(shape[0] * 100 + shape[1])
    rgbe = _random_rgbe(rng, *shape)
    cfg = api.hydrate_hdri(_hdri_config(write_hdr(tmp_path / "a.hdr", rgbe)))
    assert (cfg.hdri_width, cfg.hdri_height) == (32, 16)
    assert np.array_equal(_lut(cfg), box_lut(decode(rgbe)))


@pytest.mark.parametrize("ysign,xsign", [("+", "+"), ("-", "-"), ("+", "-")])
def test_resolution_line_signs_orient_the_image(tmp_path, ysign, xsign):
    rgbe = _random_rgbe(np.random.default_rng(7), 16, 32)
    ref = api.hydrate_hdri(_hdri_config(write_hdr(tmp_path / "a.hdr", rgbe)))
    got = api.hydrate_hdri(_hdri_config(write_hdr(tmp_path / "b.hdr", rgbe, ysign, xsign)))
    assert np.array_equal(_lut(got), _lut(ref))


def test_halfway_means_round_away_from_zero(tmp_path):
    """Texels of two pixels whose float32 mean lies exactly halfway between
    two halves (FP16 steps by 2 in [2048, 4096)): 2049 and 2053 round to 2048
    and 2052 under round-to-nearest-even, to 2050 and 2054 under the
    reference's rule; 2051 rounds to 2052 under both."""
    rgbe = np.zeros((16, 64, 4), np.uint8)
    big = (255, 136 + 4)  # 255.5 * 16 = 4088
    for row, (c, e) in enumerate([(2, 136 + 2), (4, 136 + 2), (3, 136 + 2)]):  # 10, 18, 14
        rgbe[row, 0, 0], rgbe[row, 0, 3] = big
        rgbe[row, 1, 0], rgbe[row, 1, 3] = c, e
    img = decode(rgbe)
    means = [(img[r, 0, 0] + img[r, 1, 0]) / 2 for r in range(3)]
    assert means == [2049.0, 2053.0, 2051.0]
    assert [float(np.float16(m)) for m in means] == [2048.0, 2052.0, 2052.0]  # nearest even
    lut = _lut(api.hydrate_hdri(_hdri_config(write_hdr(tmp_path / "t.hdr", rgbe))))
    got = [float(lut[r, 0, 0:1].view(np.float16)[0]) for r in range(3)]
    assert got == [2050.0, 2054.0, 2052.0]
    assert np.array_equal(lut, box_lut(img))


def test_hydration_clears_and_errors(tmp_path):
    rgbe = _random_rgbe(np.random.default_rng(1), 16, 32)
    path = write_hdr(tmp_path / "a.hdr", rgbe)
    # not world mode 2: the LUT is cleared whatever the caller put there
    c = _hdri_config(path, mode=1.0)
    for i in range(10):
        c.hdri_lut[i] = 0x3c00
    c.hdri_width = c.hdri_height = 7
    h = api.hydrate_hdri(c)
    assert h.hdri_width == 0 and h.hdri_height == 0 and not np.any(_lut(h))
    lib = ffi.load_library()

    def err_of(cfg):
        e = ffi.error_buffer()
        assert lib.fpt_hip_hydrate_hdri(C.byref(cfg), e, len(e)) == 1
        return e.value.decode()

    assert "failed to open HDRI" in err_of(_hdri_config(tmp_path / "missing.hdr"))
    assert "non-RLE" in err_of(_hdri_config(write_hdr(tmp_path / "nofmt.hdr", rgbe, fmt=False)))
    assert "non-RLE" in err_of(_hdri_config(write_hdr(tmp_path / "narrow.hdr", rgbe[:, :7])))
    blank = _hdri_config(path)
    scenes.set_hdri_path(blank, "")
    blank.hdri_enabled = 1
    assert "requires world.hdri" in err_of(blank)
    data = (tmp_path / "a.hdr").read_bytes()
    (tmp_path / "cut.hdr").write_bytes(data[: len(data) // 2])
    assert "HDRI" in err_of(_hdri_config(tmp_path / "cut.hdr"))


def test_scene_json_resolves_the_hdri_path(tmp_path):
    rgbe = _random_rgbe(np.random.default_rng(2), 16, 32)
    write_hdr(tmp_path / "sky.hdr", rgbe)
    scene = tmp_path / "s.json"
    scene.write_text('{"world": {"mode": 2, "hdri": "sky.hdr", "power": 1}}')
    c = scenes.apply_scene_json(scenes.default_config(), scene)
    assert c.hdri_enabled == 1 and bytes(c.hdri_path).rstrip(b"\0").decode() == str(tmp_path / "sky.hdr")
    assert api.hydrate_hdri(c).hdri_width == 32


# --- LUT sampling in the oracle --------------------------------------------------------

def _sampling_config(tmp_path, rng_seed=3):
    rgbe = _random_rgbe(np.random.default_rng(rng_seed), 16, 32)
    rgbe[..., 3] = np.maximum(rgbe[..., 3], 124)
    c = api.hydrate_hdri(_hdri_config(write_hdr(tmp_path / "s.hdr", rgbe)))
    c.world[4], c.world[5], c.world[2] = 1.0, 1.0, 0.0  # power 1, contrast 1: env returned as is
    return c, _lut(c).view(np.float16).astype(np.float64)


def _dir(u, v):
    th, ph = v * math.pi, 2 * math.pi * (u - 0.5)
    return [math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)]


def test_texel_centres_return_their_texel(tmp_path):
    c, lut = _sampling_config(tmp_path)
    ij = [(i, j) for j in range(1, 15) for i in range(32)]
    dirs = [_dir((i + 0.5) / 32, j / 15) for i, j in ij]
    env = bind.environment(c, dirs)
    want = np.array([lut[j, i] for i, j in ij])
    # pi is the float constant of the shading language: u is off by ~1e-8, so
    # a neighbour contributes with weight < 1e-6
    assert np.allclose(env, want, rtol=1e-5, atol=1e-6 * lut.max())


def test_seam_and_row_midpoints_average_neighbours(tmp_path):
    c, lut = _sampling_config(tmp_path)
    seam = bind.environment(c, [[-1.0, 0.0, -0.0]])[0]  # u = 0, v = 1/2
    y = 7.5  # v = 1/2 -> y = 7.5: rows 7 and 8
    want = 0.5 * (0.5 * (lut[7, 31] + lut[7, 0]) + 0.5 * (lut[8, 31] + lut[8, 0]))
    assert np.allclose(seam, want, rtol=1e-5, atol=1e-6 * lut.max()) and y == 0.5 * 15
    mids = bind.environment(c, [_dir((5 + 0.5) / 32, (j + 0.5) / 15) for j in range(1, 14)])
    assert np.allclose(mids, [0.5 * (lut[j, 5] + lut[j + 1, 5]) for j in range(1, 14)], rtol=1e-5,
                       atol=1e-6 * lut.max())


def test_yaw_of_one_column_shifts_the_lookup(tmp_path):
    c, lut = _sampling_config(tmp_path)
    c.world[2] = 360.0 / 32
    dirs = [_dir((i + 0.5) / 32, 6 / 15) for i in range(32)]
    env = bind.environment(c, dirs)
    assert np.allclose(env, [lut[6, (i - 1) % 32] for i in range(32)], rtol=1e-5, atol=1e-6 * lut.max())
    c.world[0] = 2.0
    c.hdri_enabled = 0  # mode 2 without an HDRI: constant environment 1
    assert np.allclose(bind.environment(c, dirs), 1.0)
