"""Scenes with image textures, normal maps and bump maps for the texture tests, in the reference's
.prscene.json format (textures of type "image" with a PFM / PPM / PNG "file"; a material's "reflectance",
"normal_map" and "bump_map")."""
import numpy as np

import envscenes as es


def plane_doc(tmpdir, albedo=None, normal=None, bump=None, size=16.0, film=(16, 16), fov=40.0):
    """A diffuse rect in the xz plane (u along +x, v along +z, normal +y) lit only by an overhead
    distant light, seen from above. albedo / normal: (h, w, 3) images or None; bump: (image,
    scale) or None."""
    doc = {"textures": {}, "lights": {"sun": {"type": "distant", "direction": [0, -1, 0], "irradiance": 1.0}},
           "geometry": {"plane": {"type": "rect", "size": [size, size]}},
           "objects": [{"geometry": "plane", "material": "m"}],
           "cameras": {"cam": {"type": "pinhole", "position": [0, 6, 0.5], "look_at": [0, 0, 0], "fov_y": fov}},
           "render": {"camera": "cam", "film": {"width": film[0], "height": film[1], "filter": "box"}}}
    m = {"type": "diffuse", "reflectance": 0.5}
    if albedo is not None:
        es.write_pfm(f"{tmpdir}/albedo.pfm", albedo)
        doc["textures"]["albedo"] = {"type": "image", "file": f"{tmpdir}/albedo.pfm", "wrap": "clamp"}
        m["reflectance"] = {"texture": "albedo"}
    if normal is not None:
        es.write_pfm(f"{tmpdir}/normal.pfm", normal)
        doc["textures"]["nmap"] = {"type": "image", "file": f"{tmpdir}/normal.pfm"}
        m["normal_map"] = {"texture": "nmap", "strength": 1.0}
    if bump is not None:
        img, scale = bump
        es.write_pfm(f"{tmpdir}/height.pfm", img)
        doc["textures"]["hmap"] = {"type": "image", "file": f"{tmpdir}/height.pfm", "wrap": "clamp"}
        m["bump_map"] = {"texture": "hmap", "scale": scale}
    doc["materials"] = {"m": m}
    return doc


def ramp(width=2):
    """A 1-row height image rising from 0 to 1 across its texels: with clamp wrap its bilinear
    lookup is linear in u between the first and the last texel centre."""
    v = np.linspace(0, 1, width, dtype=np.float32)
    return np.repeat(v[None, :, None], 3, axis=2)


def textured_objects_doc(tmpdir, seed=5):
    """A studio-lit set for per-sample parity: a checker-like image texture with uv tiling on the
    floor (diffuse), a normal-mapped sphere (diffuse, image albedo), a bump-mapped rough conductor
    with an image reflectance, and a sheened coated-diffuse silk with a PNG reflectance and a PNG
    weave as its bump map, under an image environment and a quad light."""
    rng = np.random.default_rng(seed)
    alb = (0.15 + 0.7 * rng.random((6, 8, 3))).astype(np.float32)
    nmap = np.clip(rng.normal([0.5, 0.5, 0.9], [0.15, 0.15, 0.05], (5, 7, 3)), 0, 1).astype(np.float32)
    height = rng.random((9, 9, 3)).astype(np.float32)
    es.write_pfm(f"{tmpdir}/t_alb.pfm", alb)
    es.write_pfm(f"{tmpdir}/t_nmap.pfm", nmap)
    es.write_pfm(f"{tmpdir}/t_height.pfm", height)
    es.write_pfm(f"{tmpdir}/t_env.pfm", es.studio_map(32, 16))
    # the silk: an 8-bit sRGB PNG reflectance and an 8-bit grey PNG weave as its bump map
    write_png(f"{tmpdir}/t_silk.png", rng.integers(120, 240, (7, 5, 3)), 2, 8, (0, 1, 4))
    write_png(f"{tmpdir}/t_weave.png", rng.integers(0, 256, (12, 12, 1)), 0, 8, (2, 3))
    return {
        "textures": {
            "alb": {"type": "image", "file": f"{tmpdir}/t_alb.pfm", "uv_scale": [3, 2], "wrap": "mirror"},
            "alb2": {"type": "image", "file": f"{tmpdir}/t_alb.pfm", "flip_v": False, "multiplier": [0.9, 0.8, 1.0]},
            "nmap": {"type": "image", "file": f"{tmpdir}/t_nmap.pfm"},
            "hmap": {"type": "image", "file": f"{tmpdir}/t_height.pfm", "uv_scale": 2, "multiplier": 0.5},
            "silk": {"type": "image", "file": f"{tmpdir}/t_silk.png"},
            "weave": {"type": "image", "file": f"{tmpdir}/t_weave.png", "colorspace": "linear", "uv_scale": 6},
        },
        "materials": {
            "floor": {"type": "diffuse", "reflectance": {"texture": "alb"}},
            "ball": {"type": "diffuse", "reflectance": {"texture": "alb2"},
                     "normal_map": {"texture": "nmap", "strength": 1.5, "flip_green": True}},
            "metal": {"type": "conductor", "reflectance": {"texture": "alb2"}, "roughness": 0.3,
                      "bump_map": {"texture": "hmap", "channel": "g", "scale": 0.02},
                      "normal_map": "nmap"},
            "silk": {"type": "coated_diffuse", "reflectance": {"texture": "silk"}, "coat_roughness": 0.28,
                     "thickness": 0.01, "sheen": {"color": [1.0, 0.96, 0.88], "roughness": 0.3, "weight": 0.8},
                     "bump_map": {"texture": "weave", "scale": 0.01}},
        },
        "geometry": {"ball": {"type": "sphere", "radius": 0.6}, "floor": {"type": "rect", "size": [6, 6]}},
        "objects": [{"geometry": "floor", "material": "floor", "transform": {"translate": [0, -0.6, 0]}},
                    {"geometry": "ball", "material": "ball", "transform": {"translate": [-0.7, 0, 0]}},
                    {"geometry": "ball", "material": "metal", "transform": {"translate": [0.7, 0, 0]}},
                    {"geometry": "ball", "material": "silk", "transform": {"translate": [0, 0.2, -1.1]}}],
        "lights": {"env": {"type": "environment", "file": f"{tmpdir}/t_env.pfm", "intensity": 0.5},
                   "key": {"type": "rect", "position": [0, 3, 2], "look_at": [0, 0, 0], "size": [1, 1],
                           "color": [1, 1, 1], "intensity": 8}},
        "cameras": {"cam": {"type": "pinhole", "position": [0, 1.2, 4], "look_at": [0, 0, 0], "fov_y": 45}},
        "render": {"camera": "cam", "film": {"width": 48, "height": 32, "filter": "box"}},
    }


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))


def write_png(path, samples, colour_type, depth, filters=(0,), interlace=False, palette=None):
    """Writes a PNG by the specification's definitions: samples (h, w, channels) integers below
    2^depth, packed MSB first below 8 bits and big-endian at 16; row r of each pass uses
    filters[r % len(filters)] (0 none, 1 sub, 2 up, 3 average, 4 Paeth); interlace = Adam7."""
    import struct
    import zlib
    h, w, nc = samples.shape
    bpp = max(1, nc * depth // 8)

    def pack(row):  # (w', nc) -> bytes of one unfiltered scanline
        flat = row.reshape(-1).astype(np.int64)
        if depth == 16:
            return np.stack([flat >> 8, flat & 255], 1).reshape(-1).astype(np.uint8)
        if depth == 8:
            return flat.astype(np.uint8)
        per = 8 // depth
        flat = np.concatenate([flat, np.zeros((-len(flat)) % per, np.int64)]).reshape(-1, per)
        shifts = 8 - depth * (np.arange(per) + 1)
        return (flat << shifts).sum(1).astype(np.uint8)

    passes = ([(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
              if interlace else [(0, 0, 1, 1)])
    out = bytearray()
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        prev = None
        for r in range(sub.shape[0]):
            raw = pack(sub[r]).astype(np.int64)
            up = prev if prev is not None else np.zeros_like(raw)
            left = np.concatenate([np.zeros(bpp, np.int64), raw[:-bpp]])
            upleft = np.concatenate([np.zeros(bpp, np.int64), up[:-bpp]])
            f = filters[r % len(filters)]
            pred = [np.zeros_like(raw), left, up, (left + up) // 2, _paeth(left, up, upleft)][f]
            out.append(f)
            out += ((raw - pred) % 256).astype(np.uint8).tobytes()
            prev = raw

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, colour_type, 0, 0, int(interlace)))
    png += chunk(b"tEXt", b"Comment\0test image")
    if palette is not None:
        png += chunk(b"PLTE", np.asarray(palette, np.uint8).tobytes())
    z = zlib.compress(bytes(out), 6)
    png += chunk(b"IDAT", z[: len(z) // 2]) + chunk(b"IDAT", z[len(z) // 2:])  # split across chunks
    png += chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


def png_expected_rgb8(samples, colour_type, depth, palette=None):
    """The 8-bit RGB a PNG's samples stand for when read as three 8-bit channels (stb_image's
    reduction): 16-bit keeps the high byte, sub-byte grey scales to 0..255, palette indices
    expand, grey replicates, alpha drops."""
    s = samples.astype(np.int64)
    if colour_type == 3:
        return np.asarray(palette, np.int64)[s[..., 0]]
    if depth == 16:
        s = s >> 8
    elif depth < 8:
        s = s * (255 // (2 ** depth - 1))
    if colour_type in (0, 4):
        return np.repeat(s[..., :1], 3, axis=2)
    return s[..., :3]


def write_hdr(path, rgb, rle=True):
    """Radiance RGBE (Ward): per texel the shared exponent e = frexp(max) + 128 and mantissas
    floor(c / 2^(e - 136)); scanlines run-length coded per component (runs of equal bytes as
    128 + n, literals as n) or flat. Returns the RGBE bytes (h, w, 4)."""
    h, w, _ = rgb.shape
    m = rgb.max(axis=2)
    mant, ex = np.frexp(m)
    e = np.where(m > 1e-32, ex + 128, 0).astype(np.int64)
    scale = np.where(e > 0, np.ldexp(1.0, -(e - 136)), 0.0)
    rgbe = np.zeros((h, w, 4), np.uint8)
    rgbe[..., :3] = np.clip(np.floor(rgb * scale[..., None]), 0, 255).astype(np.uint8)
    rgbe[..., 3] = e.astype(np.uint8)
    out = bytearray(b"#?RADIANCE\n# test map\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        if not rle or w < 8 or w > 32767:  # the format codes such scanlines flat
            out += rgbe[y].tobytes()
            continue
        out += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            row, x = rgbe[y, :, c], 0
            while x < w:
                run = 1
                while x + run < w and run < 127 and row[x + run] == row[x]:
                    run += 1
                if run >= 3:
                    out += bytes([128 + run, row[x]])
                    x += run
                else:
                    n = 1
                    while x + n < w and n < 128 and not (x + n + 2 < w and row[x + n] == row[x + n + 1] == row[x + n + 2]):
                        n += 1
                    out += bytes([n]) + row[x:x + n].tobytes()
                    x += n
    with open(path, "wb") as f:
        f.write(out)
    return rgbe


def write_exr(path, channels, compression=0, window=(0, 0), decreasing=False):
    """A single-part scanline OpenEXR file by the file layout: channels {name: (h, w) array of
    float16, float32 or uint32}; compression 0 NONE, 1 RLE, 2 ZIPS, 3 ZIP (RLE and ZIP after the
    byte predictor and the even / odd interleave); data window at `window`; decreasing: chunks
    stored bottom first."""
    import struct
    import zlib
    names = sorted(channels)
    h, w = channels[names[0]].shape
    ptype = {np.dtype(np.uint32): 0, np.dtype(np.float16): 1, np.dtype(np.float32): 2}

    def attr(name, typ, data):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<I", len(data)) + data

    chl = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", ptype[channels[n].dtype], 0, 1, 1) for n in names) + b"\0"
    x0, y0 = window
    box = struct.pack("<iiii", x0, y0, x0 + w - 1, y0 + h - 1)
    head = (attr("channels", "chlist", chl) + attr("compression", "compression", bytes([compression])) +
            attr("dataWindow", "box2i", box) + attr("displayWindow", "box2i", box) +
            attr("lineOrder", "lineOrder", bytes([1 if decreasing else 0])) +
            attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)) +
            attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0)) +
            attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    lines = 16 if compression == 3 else 1

    def predict(raw):
        t = np.frombuffer(raw, np.uint8)
        inter = np.concatenate([t[0::2], t[1::2]]).astype(np.int64)
        d = inter.copy()
        d[1:] = (inter[1:] - inter[:-1] + 128) % 256
        return d.astype(np.uint8).tobytes()

    def rle(b):
        out, i = bytearray(), 0
        while i < len(b):
            run = 1
            while i + run < len(b) and run < 128 and b[i + run] == b[i]:
                run += 1
            if run >= 3:
                out += struct.pack("<b", run - 1) + b[i:i + 1]
                i += run
            else:
                j = i
                while j < len(b) and j - i < 127 and not (j + 2 < len(b) and b[j] == b[j + 1] == b[j + 2]):
                    j += 1
                out += struct.pack("<b", -(j - i)) + b[i:j]
                i = j
        return bytes(out)

    chunks = []
    for c0 in range(0, h, lines):
        raw = b"".join(channels[n][r].astype(channels[n].dtype.newbyteorder("<")).tobytes()
                       for r in range(c0, min(h, c0 + lines)) for n in names)
        if compression == 1:
            data = rle(predict(raw))
        elif compression in (2, 3):
            data = zlib.compress(predict(raw))
        else:
            data = raw
        if len(data) >= len(raw):
            data = raw
        chunks.append(struct.pack("<ii", y0 + c0, len(data)) + data)
    order = list(range(len(chunks)))[::-1] if decreasing else list(range(len(chunks)))
    base = 8 + len(head) + 8 * len(chunks)
    offsets, body = [0] * len(chunks), b""
    for k in order:
        offsets[k] = base + len(body)
        body += chunks[k]
    with open(path, "wb") as f:
        f.write(struct.pack("<II", 20000630, 2) + head + struct.pack(f"<{len(chunks)}Q", *offsets) + body)
