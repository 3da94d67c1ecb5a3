// Image decoders for image textures and environment maps, written from the file formats'
// specifications: PNG (W3C), Radiance RGBE (.hdr) and OpenEXR (single-part scanline files).
#pragma once
#include <string>
#include <vector>

// Decodes a whole PNG file (every colour type and bit depth the PNG specification allows, plain or
// Adam7-interlaced) into 8-bit RGB, top row first: 16-bit samples keep their high byte, grey is
// replicated, palettes are expanded and alpha is dropped. Returns false with *err set on a
// malformed file.
bool DecodePNG(const std::vector<unsigned char> &file, int *w, int *h, std::vector<unsigned char> *rgb,
               std::string *err);

// Decodes a Radiance RGBE file ("-Y h +X w" layout, flat or run-length-encoded scanlines) into
// linear float RGB, top row first; a texel with exponent byte e is c * 2^(e - 136) per channel, 0
// for e = 0 (stb_image's conversion, which the reference's reader uses).
bool DecodeHDR(const std::vector<unsigned char> &file, int *w, int *h, std::vector<float> *rgb, std::string *err);

// Decodes a single-part scanline OpenEXR file (NONE, RLE, ZIPS or ZIP compression; HALF, FLOAT or
// UINT channels) into float RGB over its data window, top row first. Channels R, G and B are read
// by name; a file with a single channel is grey; a UINT channel's 32 bits are taken as a float and
// a DECREASING_Y file comes out upside down, as tinyexr's LoadEXR (the reference's reader) returns
// them. Returns false with *err set otherwise.
bool DecodeEXR(const std::vector<unsigned char> &file, int *w, int *h, std::vector<float> *rgb, std::string *err);
