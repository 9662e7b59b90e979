"""The stereo console on the GPU: the reference's own sample (Cone with its
option.xml: one level, D = 64, LR off, peak removal on), a folder of pairs
through a two-level pyramid, and the depth form.  Each written map is
compared bit for bit with the library called from Python on the same gray
bytes the console matched, and with the oracle."""
import os
import subprocess

import numpy as np
import pytest
from PIL import Image

import oracle
from fasterstereocuda_amd import scenes
from fasterstereocuda_amd.pyramid import (EpipolarStereoParam, FasterStereoCuda, StereoOption1,
                                          StereoOption2)

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONSOLE = os.path.join(ROOT, "fasterstereocuda_amd", "stereo_console")
CONE = os.path.join(ROOT, "tests", "golden", "cone")


def run(*args, cwd=None):
    r = subprocess.run([CONSOLE, *map(str, args)], cwd=cwd, capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, flush=True)
    return r


def read_pfm(path):
    data = open(path, "rb").read()
    magic, w, h, scale, rest = data.split(maxsplit=4)
    assert magic == b"Pf" and float(scale) < 0
    w, h = int(w), int(h)
    return np.frombuffer(rest, "<f4")[: w * h].reshape(h, w)[::-1].copy()


def read_pgm(path):
    data = open(path, "rb").read()
    magic, w, h, maxv, rest = data.split(maxsplit=4)
    return np.frombuffer(rest, np.uint8)[: int(w) * int(h)].reshape(int(h), int(w))


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def view8(m, left):
    valid = m != np.float32(-999999.0)
    lo, hi = m[valid].min(), m[valid].max()
    out = np.zeros(m.shape, np.uint8)
    sel = valid & (left != 0)
    out[sel] = ((m[sel] - lo) / (hi - lo) * np.float32(255)).astype(np.uint8)
    return out


def test_cone_sample_with_its_option(cuda, tmp_path):
    print("[console] Cone 450x375, its option.xml", flush=True)
    (tmp_path / "out").mkdir()
    (tmp_path / "gray").mkdir()
    r = run(os.path.join(CONE, "left.png"), os.path.join(CONE, "right.png"),
            os.path.join(CONE, "option.xml"), "--out", tmp_path / "out", "--gray", tmp_path / "gray",
            "--loops", "3")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "initialize algorithm done!" in r.stdout
    assert r.stdout.count("frame 0, timing:") == 3 and "fps =" in r.stdout
    got = read_pfm(tmp_path / "out" / "left_disp.pfm")
    L = read_pgm(tmp_path / "gray" / "left.pgm")
    R = read_pgm(tmp_path / "gray" / "right.pgm")
    assert got.shape == L.shape == (375, 450)
    # the library from Python on the same bytes and the same option
    fs = FasterStereoCuda()
    assert fs.Init(StereoOption1(width=450, height=375, min_disp=0, max_disp=64, num_layers=1,
                                 do_lr_check=False, do_rm_peaks=True)), fs.last_error
    exp = np.empty_like(got)
    assert fs.Match(L, R, exp), fs.last_error
    np.testing.assert_array_equal(bits(got), bits(exp))
    # and the oracle: 5x5 census, P1 = 8, P2 = 32, peaks at ratio 0.0005
    ref = oracle.pyramid(L, R, 0, 64, 1, 32, cs_type=0, p1=8, p2=32, do_lr_check=False,
                         peaks_ratio=0.0005)
    np.testing.assert_array_equal(bits(got), bits(ref))
    assert (got != np.float32(-999999.0)).mean() > 0.5
    png = np.asarray(Image.open(tmp_path / "out" / "left_disp.png"))
    np.testing.assert_array_equal(png, view8(got, L))


def test_folder_of_pairs_two_levels(cuda, tmp_path):
    print("[console] folder of 3 pairs 400x300, 2 levels, LR", flush=True)
    for side in ("L", "R", "out"):
        (tmp_path / side).mkdir()
    pairs = [scenes.cone_like(20 + k, H=300, W=400, D=128, workload=2) for k in range(3)]
    for k, p in enumerate(pairs):
        Image.fromarray(p.left).save(tmp_path / "L" / f"f{k}.png")
        Image.fromarray(p.right).save(tmp_path / "R" / f"f{k}.png")
    (tmp_path / "L" / "skip.txt").write_text("not an image")
    opt = tmp_path / "o.xml"
    opt.write_text("<Option1><View><width>400</width><height>300</height></View><Stereo>"
                   "<disp_min>0</disp_min><disp_max>128</disp_max><layers_num>2</layers_num>"
                   "<do_consistency_check>1</do_consistency_check>"
                   "<do_remove_peaks>0</do_remove_peaks><do_smooth>0</do_smooth></Stereo></Option1>")
    r = run(tmp_path / "L", tmp_path / "R", "png", opt, "--out", tmp_path / "out")
    assert r.returncode == 0, r.stdout + r.stderr
    fs = FasterStereoCuda()
    assert fs.Init(StereoOption1(width=400, height=300, min_disp=0, max_disp=128, num_layers=2,
                                 do_lr_check=True, do_rm_peaks=False)), fs.last_error
    for k, p in enumerate(pairs):
        got = read_pfm(tmp_path / "out" / f"f{k}_disp.pfm")
        exp = np.empty_like(got)
        assert fs.Match(p.left, p.right, exp)
        np.testing.assert_array_equal(bits(got), bits(exp), err_msg=f"frame {k}")
        ref = oracle.pyramid(p.left, p.right, 0, 128, 2, 32, cs_type=0, p1=8, p2=32)
        np.testing.assert_array_equal(bits(got), bits(ref), err_msg=f"frame {k} vs oracle")


def test_depth_form(cuda, tmp_path):
    print("[console] depth form, Option2 with a camera", flush=True)
    p = scenes.cone_like(30, H=240, W=320, D=64, workload=2)
    Image.fromarray(p.left).save(tmp_path / "l.png")
    Image.fromarray(p.right).save(tmp_path / "r.png")
    (tmp_path / "out").mkdir()
    epi = EpipolarStereoParam(x0_left=160.0, y0_left=120.0, x0_right=150.0, y0_right=120.0,
                              baseline=100.0, focus=400.0)
    opt = tmp_path / "o2.xml"
    opt.write_text(
        "<Option2><View><width>320</width><height>240</height></View><Stereo>"
        "<depth_min>500</depth_min><depth_max>4000</depth_max><layers_num>1</layers_num>"
        "<do_consistency_check>1</do_consistency_check><do_remove_peaks>0</do_remove_peaks>"
        "<do_smooth>0</do_smooth></Stereo><Camera><x0_left_pix>160</x0_left_pix>"
        "<y0_left_pix>120</y0_left_pix><x0_right_pix>150</x0_right_pix>"
        "<y0_right_pix>120</y0_right_pix><baseline_mm>100</baseline_mm>"
        "<focus_pix>400</focus_pix></Camera></Option2>")
    r = run(tmp_path / "l.png", tmp_path / "r.png", opt, "--out", tmp_path / "out")
    assert r.returncode == 0, r.stdout + r.stderr
    got = read_pfm(tmp_path / "out" / "l_disp.pfm")
    fs = FasterStereoCuda()
    assert fs.Init2(StereoOption2(width=320, height=240, min_depth=500.0, max_depth=4000.0,
                                  num_layers=1, do_lr_check=True, do_rm_peaks=False, epi=epi)), \
        fs.last_error
    exp = np.empty_like(got)
    assert fs.Match2(p.left, p.right, exp)
    np.testing.assert_array_equal(bits(got), bits(exp))
    assert (got != np.float32(-999999.0)).mean() > 0.3
