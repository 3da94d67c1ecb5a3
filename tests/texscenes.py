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
