"""File-backed datasets (host code, no GPU): the PNG decoder against PIL's
decoder and against scanlines filtered here with each PNG filter type, the
JPEG decoder against PIL's (libjpeg) within a few levels and against itself
across progressive and restart-interval codings of the same image, and
Dataset.load_transforms against the transforms.json rules of the reference
loader (src/models/dataset.cu:24-188) stated independently: the NeRF-to-this-
frame change of basis as a matrix product P M F, focal lengths from the field
of view, per-frame overrides, scene scale, missing files skipped, the most
common image size kept, and the error paths."""
import json
import struct
import zlib

import numpy as np
import pytest

PIL = pytest.importorskip("PIL.Image")


def tn():
    from turbonerf_amd import PyTurboNeRF

    return PyTurboNeRF


def _write_transforms(path, doc):
    path.write_text(json.dumps(doc))
    return str(path)


def _decode(tmp_path, files, w, h):
    """Decode image files through a file-backed Dataset (one frame each)."""
    frames = [{"file_path": f, "transform_matrix": np.eye(4).tolist()} for f in files]
    p = _write_transforms(tmp_path / "transforms.json", {"w": w, "h": h, "frames": frames})
    ds = tn().Dataset(p)
    ds.load_transforms()
    ds.load_images()
    return np.array(ds.image_data)


# ------------------------------------------------------------------ PNG
def _pil_cases(rng, w, h):
    rgba = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    rgba[0, 0, 3] = 0
    g16 = rng.integers(0, 65536, (h, w), dtype=np.uint16)
    pal = rng.integers(0, 256, (16, 3), dtype=np.uint8)
    idx = rng.integers(0, 16, (h, w), dtype=np.uint8)
    bits = rng.integers(0, 2, (h, w), dtype=np.uint8)
    cases = {}
    # image, expected RGBA8 under stb_image's conversion rules
    cases["rgba8"] = (PIL.fromarray(rgba, "RGBA"), rgba)
    cases["rgb8"] = (PIL.fromarray(rgba[..., :3], "RGB"),
                     np.concatenate([rgba[..., :3], np.full((h, w, 1), 255, np.uint8)], -1))
    g = rgba[..., 0]
    cases["grey8"] = (PIL.fromarray(g, "L"), np.stack([g, g, g, np.full_like(g, 255)], -1))
    cases["grey_alpha8"] = (PIL.fromarray(rgba[..., :2], "LA"), np.stack([g, g, g, rgba[..., 1]], -1))
    gh = (g16 >> 8).astype(np.uint8)
    cases["grey16"] = (PIL.fromarray(g16), np.stack([gh, gh, gh, np.full_like(gh, 255)], -1))
    b = (bits * 255).astype(np.uint8)
    cases["grey1"] = (PIL.fromarray(bits.astype(bool)), np.stack([b, b, b, np.full_like(b, 255)], -1))
    alpha = rng.integers(0, 256, 16, dtype=np.uint8)
    im = PIL.fromarray(idx, "P")
    im.putpalette(pal.ravel().tolist())
    im.info["transparency"] = bytes(alpha)
    cases["palette_trns"] = (im, np.concatenate([pal[idx], alpha[idx][..., None]], -1))
    key = tuple(int(v) for v in rgba[1, 1, :3])
    exp = np.concatenate([rgba[..., :3], np.full((h, w, 1), 255, np.uint8)], -1)
    exp[..., 3][np.all(rgba[..., :3] == key, -1)] = 0
    cases["rgb_key"] = (PIL.fromarray(rgba[..., :3], "RGB"), exp, key)
    return cases


def test_png_decoder_matches_pil_and_stb_rules(tmp_path):
    rng = np.random.default_rng(5)
    w, h = 37, 23
    names = []
    expected = []
    for name, case in _pil_cases(rng, w, h).items():
        fn = f"{name}.png"
        if name == "rgb_key":
            case[0].save(tmp_path / fn, transparency=case[2])
        else:
            case[0].save(tmp_path / fn)
        names.append(fn)
        expected.append(case[1])
        if name in ("rgba8", "rgb8", "grey8", "grey_alpha8", "palette_trns"):
            # PIL's own decoder agrees with the stated rule
            ref = np.array(PIL.open(tmp_path / fn).convert("RGBA"))
            assert np.array_equal(ref, case[1]), name
    got = _decode(tmp_path, names, w, h)
    for i, name in enumerate(names):
        assert np.array_equal(got[i], expected[i]), name


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)


def _png_bytes(w, h, ctype, depth, rows, filters, plte=None, interlace=0):
    """A PNG written here: each scanline filtered with the given type as
    ISO/IEC 15948 section 9 defines the filters."""
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    bpp = max(1, ch * depth // 8)
    out = bytearray()
    prior = bytes(len(rows[0]))
    for y, raw in enumerate(rows):
        ft = filters[y % len(filters)]
        f = bytearray(len(raw))
        for i in range(len(raw)):
            a = raw[i - bpp] if i >= bpp else 0
            b = prior[i]
            c = prior[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) // 2, _paeth(a, b, c)][ft]
            f[i] = (raw[i] - pred) & 0xFF
        out += bytes([ft]) + f
        prior = raw

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace))
    if plte is not None:
        png += chunk(b"PLTE", bytes(plte))
    return png + chunk(b"IDAT", zlib.compress(bytes(out), 9)) + chunk(b"IEND", b"")


def test_png_every_filter_and_bit_depth(tmp_path):
    rng = np.random.default_rng(6)
    w, h = 29, 10
    names, expected = [], []
    rgba = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    rows = [bytes(rgba[y].ravel()) for y in range(h)]
    for ft in range(5):
        fn = f"rgba_f{ft}.png"
        (tmp_path / fn).write_bytes(_png_bytes(w, h, 6, 8, rows, [ft]))
        names.append(fn)
        expected.append(rgba)
    # 16-bit RGBA, filters mixed per row: high bytes
    v16 = rng.integers(0, 65536, (h, w, 4), dtype=np.uint16)
    rows16 = [v16[y].astype(">u2").tobytes() for y in range(h)]
    (tmp_path / "rgba16.png").write_bytes(_png_bytes(w, h, 6, 16, rows16, [0, 1, 2, 3, 4]))
    names.append("rgba16.png")
    expected.append((v16 >> 8).astype(np.uint8))
    # 2- and 4-bit grey, packed from the high bit, scaled by 85 and 17
    for depth, scale in ((2, 85), (4, 17)):
        g = rng.integers(0, 1 << depth, (h, w), dtype=np.uint8)
        per = 8 // depth
        rows_g = []
        for y in range(h):
            packed = bytearray((w + per - 1) // per)
            for x in range(w):
                packed[x // per] |= int(g[y, x]) << (8 - depth * (x % per + 1))
            rows_g.append(bytes(packed))
        fn = f"grey{depth}.png"
        (tmp_path / fn).write_bytes(_png_bytes(w, h, 0, depth, rows_g, [4, 3, 1]))
        names.append(fn)
        gg = (g * scale).astype(np.uint8)
        expected.append(np.stack([gg, gg, gg, np.full_like(gg, 255)], -1))
    got = _decode(tmp_path, names, w, h)
    for i, name in enumerate(names):
        assert np.array_equal(got[i], expected[i]), name


def test_png_errors(tmp_path):
    rng = np.random.default_rng(7)
    w, h = 8, 4
    good = _png_bytes(w, h, 6, 8, [bytes(w * 4)] * h, [0])
    (tmp_path / "i.png").write_bytes(_png_bytes(w, h, 6, 8, [bytes(w * 4)] * h, [0], interlace=1))
    bad_crc = bytearray(good)
    bad_crc[30] ^= 1  # inside IHDR
    (tmp_path / "crc.png").write_bytes(bytes(bad_crc))
    (tmp_path / "trunc.png").write_bytes(good[:-20])
    PIL.fromarray(np.zeros((h + 1, w, 4), np.uint8), "RGBA").save(tmp_path / "size.png")
    for fn, msg in (("i.png", "interlaced"), ("crc.png", "CRC"), ("trunc.png", "truncated"), ("size.png", "declares")):
        with pytest.raises(RuntimeError, match=msg):
            _decode(tmp_path, [fn], w, h)


# ------------------------------------------------------------------ JPEG
def _photo(rng, h, w, c):
    """Smooth content with an edge and sensor-like noise (what JPEG is for)."""
    y, x = np.mgrid[0:h, 0:w]
    img = np.zeros((h, w, c))
    for k in range(c):
        img[..., k] = 128 + 80 * np.sin(x / (5 + 3 * k) + y / 11.0) + 30 * np.cos(y / (7.0 + k))
    img[h // 3:, w // 2:, :] += 40
    img += rng.normal(0, 6, img.shape)
    return np.clip(img, 0, 255).astype(np.uint8)


def _jpeg_cases(rng, w, h):
    rgb = _photo(rng, h, w, 3)
    grey = _photo(rng, h, w, 1)[..., 0]
    cases = {}
    for sub, tag in ((0, "444"), (1, "422"), (2, "420")):
        for q in (75, 95):
            cases[f"{tag}_q{q}"] = (PIL.fromarray(rgb), dict(quality=q, subsampling=sub))
    cases["grey"] = (PIL.fromarray(grey, "L"), dict(quality=90))
    return cases


@pytest.mark.parametrize("w,h", [(64, 48), (37, 23), (129, 67)])
def test_jpeg_decoder_matches_pil(tmp_path, w, h):
    """Baseline JPEGs written by PIL in 4:4:4, 4:2:2 and 4:2:0 at two
    qualities and in grey, sizes that are not whole MCUs: this decoder's RGBA
    against PIL's (libjpeg's integer IDCT, fixed-point colour conversion and
    triangle upsampling) within 3 levels per channel, 0.2 on average; grey
    within 1 level (the IDCT alone)."""
    rng = np.random.default_rng(w * h)
    for name, (im, kw) in _jpeg_cases(rng, w, h).items():
        fn = f"{name}.jpg"
        im.save(tmp_path / fn, format="JPEG", **kw)
        ref = np.array(PIL.open(tmp_path / fn).convert("RGBA")).astype(int)
        got = _decode(tmp_path, [fn], w, h).reshape(h, w, 4).astype(int)
        d = np.abs(got - ref)
        assert (got[..., 3] == 255).all()
        bar = 1 if name == "grey" else 3
        assert d.max() <= bar and d.mean() < 0.2, (name, d.max(), d.mean())


def test_jpeg_progressive_and_restart_decode_to_the_baseline_pixels(tmp_path):
    """Progressive files (spectral selection and successive approximation)
    and files with restart intervals hold the baseline file's quantised
    coefficients, so they decode to the same pixels bit for bit."""
    rng = np.random.default_rng(5)
    w, h = 77, 45
    im = PIL.fromarray(_photo(rng, h, w, 3))
    variants = {"base": dict(), "prog": dict(progressive=True), "rst": dict(restart_marker_blocks=3),
                "prog_rst": dict(progressive=True, restart_marker_rows=1)}
    out = {}
    for sub in (0, 2):
        for name, kw in variants.items():
            fn = f"{name}_{sub}.jpg"
            im.save(tmp_path / fn, format="JPEG", quality=90, subsampling=sub, **kw)
            out[name] = _decode(tmp_path, [fn], w, h)
        assert (tmp_path / f"prog_{sub}.jpg").read_bytes()[:4] == b"\xff\xd8\xff\xe0"
        for name in ("prog", "rst", "prog_rst"):
            assert np.array_equal(out[name], out["base"]), (sub, name)


def test_jpeg_size_from_the_frame_header(tmp_path):
    """transforms.json without w/h takes a JPEG's size from its frame
    header (after an EXIF-sized APP1 segment)."""
    rng = np.random.default_rng(6)
    im = PIL.fromarray(_photo(rng, 31, 53, 3))
    im.save(tmp_path / "a.jpg", format="JPEG", quality=90, exif=b"Exif\x00\x00" + bytes(3000))
    p = _write_transforms(tmp_path / "transforms.json",
                          {"camera_angle_x": 1.0, "frames": [{"file_path": "a.jpg", "transform_matrix": np.eye(4).tolist()}]})
    ds = tn().Dataset(p)
    ds.load_transforms()
    assert tuple(ds.image_dimensions) == (53, 31)
    ds.load_images()
    assert np.array(ds.image_data).size == 53 * 31 * 4


def test_jpeg_errors(tmp_path):
    rng = np.random.default_rng(8)
    w, h = 16, 16
    PIL.fromarray(_photo(rng, h, w, 3)).save(tmp_path / "ok.jpg", format="JPEG", quality=90)
    good = (tmp_path / "ok.jpg").read_bytes()
    sof = good.index(b"\xff\xc0")
    arith = bytearray(good)
    arith[sof + 1] = 0xC9  # arithmetic-coded frame
    (tmp_path / "arith.jpg").write_bytes(bytes(arith))
    (tmp_path / "trunc.jpg").write_bytes(good[:sof + 6])
    twelve = bytearray(good)
    twelve[sof + 4] = 12  # sample precision
    (tmp_path / "p12.jpg").write_bytes(bytes(twelve))
    PIL.fromarray(_photo(rng, h, w, 4), "CMYK").save(tmp_path / "cmyk.jpg", format="JPEG")
    for fn, msg in (("arith.jpg", "arithmetic"), ("trunc.jpg", "truncated"), ("p12.jpg", "12-bit"),
                    ("cmyk.jpg", "4 components")):
        with pytest.raises(RuntimeError, match=msg):
            _decode(tmp_path, [fn], w, h)


# ------------------------------------------------------------ transforms.json
P = np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]], np.float32)  # rows xyz <- yzx
F = np.diag([1, -1, -1, 1]).astype(np.float32)               # flip the y and z columns


def _expected_transform(m44, scale):
    return (np.float32(scale) * (P @ (np.asarray(m44, np.float32)[:3] @ F))).astype(np.float32)


def test_load_transforms_rules(tmp_path):
    t = tn()
    rng = np.random.default_rng(8)
    for name, (w, h) in (("a", (40, 30)), ("b", (40, 30)), ("c", (40, 30)), ("d", (20, 10)), ("e", (20, 10))):
        PIL.fromarray(rng.integers(0, 256, (h, w, 4), dtype=np.uint8), "RGBA").save(tmp_path / f"{name}.png")
    mats = [rng.normal(size=(4, 4)).round(3).tolist() for _ in range(6)]
    doc = {
        "camera_angle_x": 0.9, "camera_angle_y": 0.7, "w": 40, "h": 30, "aabb_scale": 256, "scene_scale": 0.5,
        "near": 0.1, "far": 20.0, "k1": -0.1, "p2": 0.002,
        "frames": [
            {"file_path": "a.png", "transform_matrix": mats[0]},
            {"file_path": "b.png", "transform_matrix": mats[1], "fl_x": 33.5, "cx": 19.0, "k1": 0.05, "near": 0.2},
            {"file_path": "gone.png", "transform_matrix": mats[2]},
            {"file_path": "c.png", "transform_matrix": mats[3]},
            {"file_path": "d.png", "transform_matrix": mats[4], "w": 20, "h": 10},
            {"file_path": "e.png", "transform_matrix": mats[5], "w": 20, "h": 10},
        ]}
    ds = t.Dataset(_write_transforms(tmp_path / "transforms.json", doc))
    ds.load_transforms()
    # the missing file is skipped; 40x30 (three frames) outnumbers 20x10 (two)
    assert ds.n_images == 3 and ds.contains_multiple_image_dims and ds.image_dimensions == (40, 30)
    assert [p.rsplit("/", 1)[-1] for p in ds.image_paths] == ["a.png", "b.png", "c.png"]
    assert ds.bounding_box.max_x == 64.0 and ds.bounding_box.min_z == -64.0  # aabb_scale capped at 128
    fx = np.float32(0.5) * np.float32(40) / np.tan(np.float32(0.5) * np.float32(0.9), dtype=np.float32)
    fy = np.float32(0.5) * np.float32(30) / np.tan(np.float32(0.5) * np.float32(0.7), dtype=np.float32)
    for cam, m, over in zip(ds.cameras, (mats[0], mats[1], mats[3]), (False, True, False)):
        np.testing.assert_array_equal(cam.transform.to_matrix(), _expected_transform(m, 0.5))
        np.testing.assert_allclose(cam.focal_length, (33.5 if over else fx, fy), rtol=1e-6)
        assert cam.principal_point == ((19.0 if over else 20.0), 15.0)
        assert cam.near == np.float32(0.5) * np.float32(0.2 if over else 0.1)
        assert cam.far == np.float32(10.0)
        d = cam.dist_params
        assert (d.k1, d.k2, d.p2) == (np.float32(0.05 if over else -0.1), 0.0, np.float32(0.002))
        assert cam.resolution == (40, 30)


def test_load_transforms_tie_fl_defaults_and_errors(tmp_path):
    t = tn()
    rng = np.random.default_rng(9)
    for name, (w, h) in (("a", (12, 8)), ("b", (10, 9))):
        PIL.fromarray(rng.integers(0, 256, (h, w, 4), dtype=np.uint8), "RGBA").save(tmp_path / f"{name}.png")
    # a tie between sizes keeps the smaller width; no w/h given: sizes from the PNG headers;
    # no field of view: fl_x defaults to 1000 and fl_y to fl_x
    doc = {"fl_x": 500.0, "frames": [{"file_path": "a.png", "transform_matrix": np.eye(4).tolist()},
                                     {"file_path": "b.png", "transform_matrix": np.eye(4).tolist()}]}
    ds = t.Dataset(_write_transforms(tmp_path / "t1.json", doc))
    ds.load_transforms()
    assert ds.n_images == 1 and ds.image_dimensions == (10, 9)
    assert ds.cameras[0].focal_length == (500.0, 500.0) and ds.cameras[0].principal_point == (0.0, 0.0)
    assert ds.bounding_box.max_x == 8.0  # aabb_scale defaults to 16
    doc2 = {"frames": [{"file_path": "a.png", "transform_matrix": np.eye(4).tolist()}]}
    ds = t.Dataset(_write_transforms(tmp_path / "t2.json", doc2))
    ds.load_transforms()
    assert ds.cameras[0].focal_length == (1000.0, 1000.0)
    with pytest.raises(RuntimeError, match="No valid images"):
        d = t.Dataset(_write_transforms(tmp_path / "t3.json", {"frames": [{"file_path": "no.png",
                                                                            "transform_matrix": []}]}))
        d.load_transforms()
    (tmp_path / "bad.json").write_text('{"frames": [1, 2,, 3]}')
    with pytest.raises(RuntimeError, match="byte"):
        t.Dataset(str(tmp_path / "bad.json")).load_transforms()
    with pytest.raises(RuntimeError, match="Could not open"):
        t.Dataset(str(tmp_path / "absent.json")).load_transforms()
    (tmp_path / "t4.json").write_text(json.dumps({"w": "wide", "frames": []}))
    with pytest.raises(RuntimeError, match="not a number"):
        t.Dataset(str(tmp_path / "t4.json")).load_transforms()


def test_round_trip_of_this_frames_cameras(tmp_path):
    """Cameras written as NeRF matrices (the inverse change of basis,
    P^T M' F) and read back are the same float32 transforms."""
    from turbonerf_amd import scenes

    t = tn()
    cams = scenes.sphere_cameras(10, 5, 2.0, 16, 12, 50.0)
    PIL.fromarray(np.zeros((12, 16, 4), np.uint8), "RGBA").save(tmp_path / "z.png")
    frames = []
    for c in cams:
        m = np.array(c.transform[:], np.float32).reshape(3, 4)
        nerf = np.eye(4, dtype=np.float32)
        nerf[:3] = P.T @ m @ F
        frames.append({"file_path": "z.png", "transform_matrix": nerf.astype(np.float64).tolist(),
                       "fl_x": float(c.focal_length[0]), "fl_y": float(c.focal_length[1])})
    doc = {"w": 16, "h": 12, "near": float(cams[0].near_plane), "far": float(cams[0].far_plane), "frames": frames}
    ds = t.Dataset(_write_transforms(tmp_path / "transforms.json", doc))
    ds.load_transforms()
    for c, got in zip(cams, ds.cameras):
        np.testing.assert_array_equal(got.transform.to_matrix(), np.array(c.transform[:], np.float32).reshape(3, 4))
        assert got.near == np.float32(c.near_plane) and got.far == np.float32(c.far_plane)
        assert got.principal_point == (8.0, 6.0)


def test_to_json_round_trip_and_images(tmp_path):
    """Dataset.to_json writes NeRF-convention matrices, posix paths relative
    to the dataset file and scene scale 1; reading it back gives the same
    cameras, paths and box (scene scale and per-frame overrides already
    applied). Dataset.images lists the image files with their size."""
    t = tn()
    rng = np.random.default_rng(11)
    (tmp_path / "imgs").mkdir()
    for name in ("a", "b"):
        PIL.fromarray(rng.integers(0, 256, (9, 14, 4), dtype=np.uint8), "RGBA").save(tmp_path / "imgs" / f"{name}.png")
    doc = {"camera_angle_x": 0.8, "w": 14, "h": 9, "aabb_scale": 8, "scene_scale": 0.3, "k2": 0.01,
           "frames": [{"file_path": "imgs/a.png", "transform_matrix": rng.normal(size=(4, 4)).tolist()},
                      {"file_path": "imgs/b.png", "transform_matrix": rng.normal(size=(4, 4)).tolist(),
                       "fl_y": 21.5, "p1": 0.003}]}
    ds = t.Dataset(_write_transforms(tmp_path / "transforms.json", doc))
    ds.load_transforms()
    assert [(im.file_path.rsplit("/", 2)[-2:], im.dimensions) for im in ds.images] == \
        [(["imgs", "a.png"], (14, 9)), (["imgs", "b.png"], (14, 9))]
    out = ds.to_json()
    assert out["scene_scale"] == 1.0 and out["aabb_scale"] == 8 and (out["w"], out["h"]) == (14, 9)
    assert [f["file_path"] for f in out["frames"]] == ["imgs/a.png", "imgs/b.png"]
    ds2 = t.Dataset(_write_transforms(tmp_path / "again.json", out))
    ds2.load_transforms()
    assert ds2.image_paths == ds.image_paths and ds2.bounding_box == ds.bounding_box
    for a, b in zip(ds.cameras, ds2.cameras):
        np.testing.assert_array_equal(a.transform.to_matrix(), b.transform.to_matrix())
        assert (a.focal_length, a.principal_point, a.near, a.far, a.resolution) == \
            (b.focal_length, b.principal_point, b.near, b.far, b.resolution)
        da, db = a.dist_params, b.dist_params
        assert (da.k1, da.k2, da.k3, da.p1, da.p2) == (db.k1, db.k2, db.k3, db.p1, db.p2)
    with pytest.raises(RuntimeError, match="to_json"):
        t.Dataset(t.BoundingBox(1.0), list(ds.cameras), np.zeros((2, 9, 14, 4), np.uint8)).to_json()


def test_render_buffer_save_image(tmp_path):
    """RenderTarget.save_image writes the float RGBA buffer as an 8-bit RGBA
    PNG with (uint8)(v * 255), truncated (reference gpu-image.cu:13-24), values
    outside [0, 1] clamped; PIL reads back exactly those bytes."""
    t = tn()
    buf = t.CPURenderBuffer()
    buf.set_size(13, 7)
    rgba = np.asarray(buf.get_rgba())
    rng = np.random.default_rng(12)
    vals = rng.uniform(-0.2, 1.2, size=rgba.shape).astype(np.float32)
    vals[0, 0] = [0.999, 1.0 / 255.0, 0.5, 1.0]
    rgba[...] = vals
    buf.save_image(str(tmp_path / "out.png"))
    got = np.array(PIL.open(tmp_path / "out.png"))
    want = np.clip(np.trunc(vals * np.float32(255)), 0, 255).astype(np.uint8)
    assert got.shape == (7, 13, 4) and np.array_equal(got, want)
    buf.free()
    with pytest.raises(RuntimeError, match="no image"):
        buf.save_image(str(tmp_path / "none.png"))


def test_transforms_json_escapes_and_number_forms(tmp_path):
    """File names written with \\u escapes (a surrogate pair among them) and
    matrices written with exponents, signs and integers read back as the
    names and values they denote."""
    t = tn()
    name = "béb\U0001F600.png"
    PIL.fromarray(np.zeros((4, 6, 4), np.uint8), "RGBA").save(tmp_path / name)
    m = [[1e0, -0.0, 0, 2.5E-1], [0, 1, 0, -3e+2], [0, 0, 1.0, 1e-3], [0, 0, 0, 1]]
    text = ('{"w": 6, "h": 4, "frames": [ {"file_path": "b\\u00e9b\\ud83d\\ude00.png", '
            '"transform_matrix": [[1e0, -0.0, 0, 2.5E-1], [0, 1, 0, -3e+2], [0, 0, 1.0, 1e-3], [0, 0, 0, 1]]} ]}')
    (tmp_path / "t.json").write_text(text, encoding="utf-8")
    ds = t.Dataset(str(tmp_path / "t.json"))
    ds.load_transforms()
    assert ds.image_paths[0].endswith(name)
    P = np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]], np.float32)
    F = np.diag([1, -1, -1, 1]).astype(np.float32)
    want = P @ (np.array(m, np.float32)[:3] @ F)
    np.testing.assert_array_equal(ds.cameras[0].transform.to_matrix(), want)
    for bad in ('{"frames": [}', '{"frames": []} x', '{"a": "\\ud83d"}', '{"a": 01}', '{"a": 1.}'):
        (tmp_path / "bad.json").write_text(bad)
        with pytest.raises(RuntimeError, match="JSON"):
            t.Dataset(str(tmp_path / "bad.json")).load_transforms()
