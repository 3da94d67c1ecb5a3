// PNG decoding for image textures and environment maps.
#pragma once
#include <string>
#include <vector>

// Decodes a whole PNG file (every colour type and bit depth the PNG specification allows, plain or
// Adam7-interlaced) into 8-bit RGB, top row first: 16-bit samples keep their high byte, grey is
// replicated, palettes are expanded and alpha is dropped. Returns false with *err set on a
// malformed file.
bool DecodePNG(const std::vector<unsigned char> &file, int *w, int *h, std::vector<unsigned char> *rgb,
               std::string *err);
