"""fpt_compare_images (metal_bridge.mm:2358-2497, 6062-6145): host-side image
metrics, pinned by closed forms (identity, constant offsets, constant-block
SSIM), gate selection by path, the PNG decoder over every colour type and row
filter, and the report / comparison-sheet outputs.  No GPU work."""
import ctypes as C
import json
import struct
import zlib

import numpy as np
import pytest

from fractalpathtracer_amd import ffi

C1, C2 = (0.01 * 255) ** 2, (0.03 * 255) ** 2


def _png(path, img, ctype=6, filt=0):
    """Minimal PNG encoder: 8-bit, colour type 0/2/4/6, one filter type for
    every row (0 none, 1 sub, 2 up, 3 average, 4 Paeth)."""
    img = np.asarray(img, np.uint8)
    h, w = img.shape[:2]
    ch = {0: 1, 4: 2, 2: 3, 6: 4}[ctype]
    px = img.reshape(h, w, -1)[..., :ch].reshape(h, w * ch).astype(np.int32)
    raw = bytearray()
    prev = np.zeros(w * ch, np.int32)
    for y in range(h):
        cur = px[y]
        out = np.zeros_like(cur)
        for x in range(w * ch):
            a = cur[x - ch] if x >= ch else 0
            b = prev[x]
            c = prev[x - ch] if x >= ch else 0
            if filt == 0:
                pred = 0
            elif filt == 1:
                pred = a
            elif filt == 2:
                pred = b
            elif filt == 3:
                pred = (a + b) // 2
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            out[x] = (cur[x] - pred) & 255
        raw += bytes([filt]) + bytes(out.astype(np.uint8))
        prev = cur

    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)

    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0))
    data += chunk(b"IDAT", zlib.compress(bytes(raw))) + chunk(b"IEND", b"")
    path.write_bytes(data)
    return str(path)


def _compare(tmp_path, a, b, name="cand.png", **kw):
    lib = ffi.load_library()
    pa = _png(tmp_path / "base.png", a, **kw)
    pb = _png(tmp_path / name, b, **kw)
    rep = tmp_path / "report.json"
    err = ffi.error_buffer()
    rc = lib.fpt_compare_images(pa.encode(), pb.encode(), str(rep).encode(), err, len(err))
    if rc != 0:
        raise ffi.FptError(err.value.decode())
    return json.loads(rep.read_text())


def _rgba(h, w, rgb):
    img = np.zeros((h, w, 4), np.uint8)
    img[..., :3] = rgb
    img[..., 3] = 255
    return img


def test_identical_images(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.integers(0, 256, (37, 53, 4), dtype=np.uint8)
    r = _compare(tmp_path, a, a)
    assert r["mean_absolute_error"] == 0 and r["root_mean_square_error"] == 0 and r["max_absolute_error"] == 0
    assert r["luminance_ssim"] == 1.0 and r["low_frequency_luminance_ssim"] == 1.0
    assert r["changed_pixel_percent"] == 0 and r["width"] == 53 and r["height"] == 37
    assert r["unique_colours"]["candidate"] > 64 and r["recognition_gate"] and r["strict_gate"]


def test_constant_offset_closed_forms(tmp_path):
    a = _rgba(16, 24, [100, 50, 20])
    b = _rgba(16, 24, [107, 57, 27])
    r = _compare(tmp_path, a, b)
    assert r["mean_absolute_error"] == pytest.approx(7) and r["root_mean_square_error"] == pytest.approx(7)
    assert r["max_absolute_error"] == 7 and r["changed_pixel_percent"] == pytest.approx(100)
    la = 0.2126 * 100 + 0.7152 * 50 + 0.0722 * 20
    lb = la + 7
    want = (2 * la * lb + C1) / (la * la + lb * lb + C1)  # zero variance and covariance
    assert r["luminance_ssim"] == pytest.approx(want, rel=1e-12)
    assert r["low_frequency_luminance_ssim"] == pytest.approx(want, rel=1e-12)
    # 1 colour each: below the recognition gate's 16-colour floor
    assert r["unique_colours"] == {"baseline": 1, "candidate": 1} and not r["recognition_gate"]


def test_block_ssim_weights_partial_blocks_by_pixel_count(tmp_path):
    """Left 8 columns equal, right 2 columns offset: blocks are 8x8 and 8x2,
    weighted by their pixel counts (64 and 16)."""
    a = _rgba(8, 10, [80, 80, 80])
    b = a.copy()
    b[:, 8:, :3] = 120
    r = _compare(tmp_path, a, b)
    ssim_edge = (2 * 80 * 120 + C1) / (80 ** 2 + 120 ** 2 + C1)
    assert r["luminance_ssim"] == pytest.approx((64 * 1.0 + 16 * ssim_edge) / 80, rel=1e-12)
    assert r["changed_pixel_percent"] == pytest.approx(20)


@pytest.mark.parametrize("ctype", [0, 2, 4, 6])
@pytest.mark.parametrize("filt", [0, 1, 2, 3, 4])
def test_png_decoder_colour_types_and_filters(tmp_path, ctype, filt):
    rng = np.random.default_rng(ctype * 10 + filt)
    a = rng.integers(0, 256, (9, 13, 4), dtype=np.uint8)
    b = a.copy()
    b[2, 3, :] = (b[2, 3, :].astype(int) + 40) % 256
    if ctype in (0, 4):   # grey: the decoder replicates the grey value into RGB
        a[..., 1] = a[..., 2] = a[..., 0]
        b[..., 1] = b[..., 2] = b[..., 0]
    ref = _compare(tmp_path, a, b, ctype=6, filt=0)
    got = _compare(tmp_path, a, b, ctype=ctype, filt=filt)
    for k in ("mean_absolute_error", "root_mean_square_error", "luminance_ssim", "low_frequency_luminance_ssim",
              "max_absolute_error", "changed_pixel_percent"):
        assert got[k] == pytest.approx(ref[k], rel=1e-12), k


@pytest.mark.parametrize("name,gate", [("x/Mandelbox_Fractal.png", "mandelbox_fractal"),
                                       ("preview-parity/Glass_Ball.png", "preview_glass"),
                                       ("a/Cornell_Box-hip.png", "cornell"), ("other.png", "generic")])
def test_gate_class_from_candidate_path(tmp_path, name, gate):
    (tmp_path / name).parent.mkdir(parents=True, exist_ok=True)
    a = _rgba(8, 8, [10, 20, 30])
    r = _compare(tmp_path, a, a, name=name)
    assert r["gate_class"] == gate


def test_report_sheet_and_errors(tmp_path):
    a = _rgba(20, 40, [10, 20, 30])
    b = _rgba(20, 40, [10, 20, 90])
    r = _compare(tmp_path, a, b)
    sheet = tmp_path / "report.comparison.png"
    assert r["comparison_sheet"] == str(sheet) and sheet.exists()
    w, h = struct.unpack(">II", sheet.read_bytes()[16:24])
    assert (w, h) == (360 * 3 + 24, 180)
    with pytest.raises(ffi.FptError, match="dimension mismatch"):
        _compare(tmp_path, a, _rgba(21, 40, [0, 0, 0]))
    lib = ffi.load_library()
    err = ffi.error_buffer()
    assert lib.fpt_compare_images(b"/nonexistent.png", b"/nonexistent.png", str(tmp_path / "r.json").encode(),
                                  err, len(err)) == 1
    assert b"failed to read image" in err.value
