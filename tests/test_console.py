"""The stereo console's host side on the CPU: option.xml reading and writing
(the schema of FasterStereoConsole/option_manager.cpp:46-340), the argument
modes that need no GPU, and the image decoder against independent encoders
(PIL, and PNG streams built here with every filter type and Adam7)."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest
from PIL import Image

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONSOLE = os.path.join(ROOT, "fasterstereocuda_amd", "stereo_console")
CONE = os.path.join(ROOT, "tests", "golden", "cone")

pytestmark = pytest.mark.skipif(not os.path.exists(CONSOLE), reason="stereo_console is not built")


def run(*args, cwd=None):
    return subprocess.run([CONSOLE, *map(str, args)], cwd=cwd, capture_output=True, text=True,
                          timeout=60)


def show(path):
    r = run("--show-option", path)
    assert r.returncode == 0, r.stdout + r.stderr
    return dict(line.split("=", 1) for line in r.stdout.split())


def luma(rgb):
    rgb = rgb.astype(np.uint32)
    return ((9797 * rgb[..., 0] + 19234 * rgb[..., 1] + 3737 * rgb[..., 2] + 16384) >> 15).astype(np.uint8)


def read_pgm(path):
    data = open(path, "rb").read()
    magic, w, h, maxv, rest = data.split(maxsplit=4)
    assert magic == b"P5" and maxv == b"255"
    return np.frombuffer(rest, np.uint8)[: int(w) * int(h)].reshape(int(h), int(w))


def to_gray(tmp_path, image):
    out = tmp_path / "g.pgm"
    r = run("--to-gray", image, out)
    assert r.returncode == 0, r.stdout
    return read_pgm(out)


# ------------------------------------------------------------ option files
def test_reads_the_reference_cone_option():
    o = show(os.path.join(CONE, "option.xml"))
    assert o == {"form": "Option1", "width": "450", "height": "375", "min_disp": "0",
                 "max_disp": "64", "num_layers": "1", "do_lr_check": "0", "do_rm_peaks": "1",
                 "do_smooth": "0"}


def test_reads_depth_option_with_camera(tmp_path):
    p = tmp_path / "o2.xml"
    p.write_text("""<?xml version="1.0"?>
<!-- a comment before the root -->
<Option2 version="2">
  <View><width> 640 </width><height>480</height></View>
  <Stereo>
    <depth_min>500.25</depth_min><depth_max>4e3</depth_max>
    <layers_num>3</layers_num><do_consistency_check>1</do_consistency_check>
    <do_remove_peaks>0</do_remove_peaks><do_smooth>0</do_smooth>
  </Stereo>
  <Camera>
    <x0_left_pix>320.5</x0_left_pix><y0_left_pix>240</y0_left_pix>
    <x0_right_pix>330.75</x0_right_pix><y0_right_pix>241</y0_right_pix>
    <baseline_mm>120.125</baseline_mm><focus_pix>700.0</focus_pix>
  </Camera>
</Option2>
""")
    o = show(p)
    assert o["form"] == "Option2"
    assert (o["width"], o["height"], o["num_layers"]) == ("640", "480", "3")
    assert (o["do_lr_check"], o["do_rm_peaks"], o["do_smooth"]) == ("1", "0", "0")
    assert float(o["min_depth"]) == np.float32(500.25) and float(o["max_depth"]) == 4000.0
    assert float(o["x0_left"]) == 320.5 and float(o["x0_right"]) == 330.75
    assert float(o["y0_left"]) == 240.0 and float(o["y0_right"]) == 241.0
    assert float(o["baseline"]) == 120.125 and float(o["focus"]) == 700.0


def test_values_read_as_atoi_and_missing_elements_keep_defaults(tmp_path):
    p = tmp_path / "o.xml"
    # atoi: leading blanks skipped, trailing text ignored, no digits -> 0;
    # entities and CDATA are character data; unknown elements are ignored
    p.write_text("<Option1><View><width>\n 12px</width><extra a='1'/></View>"
                 "<Stereo><disp_max><![CDATA[1]]>28</disp_max><disp_min>&#45;16</disp_min>"
                 "<do_smooth>yes</do_smooth></Stereo></Option1>")
    o = show(p)
    assert o["width"] == "12" and o["height"] == "0"
    assert o["max_disp"] == "128" and o["min_disp"] == "-16"
    assert o["do_smooth"] == "0"
    # defaults of StereoOption1 where the file has no element
    assert (o["num_layers"], o["do_lr_check"], o["do_rm_peaks"]) == ("1", "1", "1")


def test_opt_mode_writes_both_default_files(tmp_path):
    r = run("opt", cwd=tmp_path)
    assert r.returncode == 0
    assert "initialize parameter file of algorithm done!" in r.stdout
    text = (tmp_path / "option.xml").read_text()
    assert text == (
        '<?xml version="1.0" encoding="UTF-8" standalone="no"?>\n<Option1>\n    <View>\n'
        "        <width>0</width>\n        <height>0</height>\n    </View>\n    <Stereo>\n"
        "        <disp_min>0</disp_min>\n        <disp_max>0</disp_max>\n"
        "        <layers_num>1</layers_num>\n        <do_consistency_check>1</do_consistency_check>\n"
        "        <do_remove_peaks>1</do_remove_peaks>\n        <do_smooth>0</do_smooth>\n"
        "    </Stereo>\n</Option1>\n")
    o2 = show(tmp_path / "option2.xml")
    assert o2["form"] == "Option2" and o2["min_depth"] == "0" and o2["focus"] == "0"
    # existing files are kept
    (tmp_path / "option.xml").write_text("<Option1><View><width>7</width></View></Option1>")
    run("opt", cwd=tmp_path)
    assert show(tmp_path / "option.xml")["width"] == "7"


@pytest.mark.parametrize("doc", ["<Option1><View></Option1>", "<Option1>", "no xml at all",
                                 "<Option1><a>&bogus;</a></Option1>", "<Option1/><Option2/>"])
def test_malformed_option_is_refused_and_kept(tmp_path, doc):
    p = tmp_path / "bad.xml"
    p.write_text(doc)
    r = run("l.png", "r.png", p)
    assert r.returncode == 1
    assert "loading parameters of algorithm failed!" in r.stdout
    assert p.read_text() == doc


def test_unknown_root_and_missing_option_file(tmp_path):
    p = tmp_path / "x.xml"
    p.write_text("<Option3><View/></Option3>")
    assert run("l.png", "r.png", p).returncode == 1
    assert run("l.png", "r.png", tmp_path / "missing.xml").returncode == 1


def test_argument_errors():
    assert run().returncode == 2
    assert run("a", "b").returncode == 2
    assert run("a", "b", "c", "d", "e").returncode == 2
    assert run("a", "b", "c", "--bogus", "1").returncode == 2
    assert run("a", "b", "c", "--loops", "0").returncode == 2


def test_image_size_must_match_the_option(tmp_path):
    Image.fromarray(np.zeros((20, 30), np.uint8)).save(tmp_path / "l.png")
    Image.fromarray(np.zeros((20, 31), np.uint8)).save(tmp_path / "r.png")
    r = run(tmp_path / "l.png", tmp_path / "r.png", os.path.join(CONE, "option.xml"))
    assert r.returncode == 1 and "data error!" in r.stdout


def test_folder_mode_needs_equal_counts(tmp_path):
    (tmp_path / "L").mkdir()
    (tmp_path / "R").mkdir()
    for k in range(2):
        Image.fromarray(np.zeros((20, 30), np.uint8)).save(tmp_path / "L" / f"{k}.png")
    Image.fromarray(np.zeros((20, 30), np.uint8)).save(tmp_path / "R" / "0.png")
    r = run(tmp_path / "L", tmp_path / "R", "png", os.path.join(CONE, "option.xml"))
    assert r.returncode == 1 and "2 left and 1 right" in r.stdout


# ------------------------------------------------------------ image decoding
def test_cone_png_decodes_to_bt601_gray(tmp_path):
    g = to_gray(tmp_path, os.path.join(CONE, "left.png"))
    rgb = np.asarray(Image.open(os.path.join(CONE, "left.png")).convert("RGB"))
    np.testing.assert_array_equal(g, luma(rgb))


@pytest.mark.parametrize("mode", ["L", "RGB", "RGBA", "LA", "P", "1", "I;16"])
def test_png_colour_types_against_pil(tmp_path, mode):
    rng = np.random.default_rng(len(mode))
    H, W = 37, 53
    rgb = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    if mode == "I;16":
        v = rng.integers(0, 65536, (H, W), dtype=np.uint16)
        im = Image.fromarray(v)
        assert im.mode == "I;16"
        exp = (v >> 8).astype(np.uint8)
    elif mode == "1":
        b = rng.random((H, W)) < 0.5
        im = Image.fromarray(b)
        exp = np.where(np.asarray(im), 255, 0).astype(np.uint8)
    elif mode == "P":
        im = Image.fromarray(rgb).quantize(colors=23)
        exp = luma(np.asarray(im.convert("RGB")))
    elif mode in ("L", "LA"):
        im = Image.fromarray(rgb[..., 0]).convert(mode)
        exp = rgb[..., 0]
    else:
        im = Image.fromarray(rgb).convert(mode)
        exp = luma(rgb)
    p = tmp_path / "i.png"
    im.save(p)
    np.testing.assert_array_equal(to_gray(tmp_path, p), exp)


def _chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data))


def _filter_row(kind, row, prev, bpp):
    """PNG filter `kind` of one row (the encoder side of the five filters)."""
    row = row.astype(np.int32)
    prev = prev.astype(np.int32)
    left = np.concatenate([np.zeros(bpp, np.int32), row[:-bpp]])
    upleft = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]])
    if kind == 0:
        pred = np.zeros_like(row)
    elif kind == 1:
        pred = left
    elif kind == 2:
        pred = prev
    elif kind == 3:
        pred = (left + prev) // 2
    else:
        p = left + prev - upleft
        pa, pb, pc = np.abs(p - left), np.abs(p - prev), np.abs(p - upleft)
        pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
    return ((row - pred) & 0xFF).astype(np.uint8)


def _encode_rgb(img, filters, interlace):
    H, W, _ = img.shape
    passes = ([(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
               (0, 1, 1, 2)] if interlace else [(0, 0, 1, 1)])
    raw = bytearray()
    k = 0
    for x0, y0, dx, dy in passes:
        sub = img[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        prev = np.zeros(sub.shape[1] * 3, np.uint8)
        for row in sub.reshape(sub.shape[0], -1):
            f = filters[k % len(filters)]
            k += 1
            raw.append(f)
            raw += _filter_row(f, row, prev, 3).tobytes()
            prev = row
    ihdr = struct.pack(">IIBBBBB", W, H, 8, 2, 0, 0, 1 if interlace else 0)
    idat = zlib.compress(bytes(raw), 9)
    # split the data over two IDAT chunks, as encoders may
    return (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"tEXt", b"k\x00v")
            + _chunk(b"IDAT", idat[:7]) + _chunk(b"IDAT", idat[7:]) + _chunk(b"IEND", b""))


@pytest.mark.parametrize("interlace", [False, True], ids=["progressive", "adam7"])
def test_png_filters_and_interlace(tmp_path, interlace):
    rng = np.random.default_rng(5)
    H, W = 29, 41
    img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    img[:, ::2] //= 3                 # structure, so the predictors differ
    p = tmp_path / "f.png"
    p.write_bytes(_encode_rgb(img, [0, 1, 2, 3, 4], interlace))
    # the stream is valid PNG: PIL reads back the same pixels
    np.testing.assert_array_equal(np.asarray(Image.open(p).convert("RGB")), img)
    np.testing.assert_array_equal(to_gray(tmp_path, p), luma(img))


def test_corrupt_png_is_refused(tmp_path):
    rng = np.random.default_rng(6)
    data = bytearray(_encode_rgb(rng.integers(0, 256, (8, 8, 3), dtype=np.uint8), [0], False))
    data[40] ^= 0xFF                  # inside IHDR / IDAT: the CRC no longer matches
    p = tmp_path / "c.png"
    p.write_bytes(bytes(data))
    r = run("--to-gray", p, tmp_path / "o.pgm")
    assert r.returncode == 1 and "PNG" in r.stdout


def test_pgm_binary_and_ascii(tmp_path):
    rng = np.random.default_rng(7)
    img = rng.integers(0, 256, (9, 13), dtype=np.uint8)
    p5 = tmp_path / "a.pgm"
    p5.write_bytes(b"P5\n# comment\n13 9\n255\n" + img.tobytes())
    np.testing.assert_array_equal(to_gray(tmp_path, p5), img)
    p2 = tmp_path / "b.pgm"
    p2.write_text("P2\n13 9\n255\n" + "\n".join(" ".join(map(str, r)) for r in img) + "\n")
    np.testing.assert_array_equal(to_gray(tmp_path, p2), img)
    p16 = tmp_path / "c.pgm"
    v = img.astype(np.uint16) * 257
    p16.write_bytes(b"P5 13 9 65535\n" + v.astype(">u2").tobytes())
    np.testing.assert_array_equal(to_gray(tmp_path, p16), img)
