// png.hpp -- PNG -> RGBA8 with a vertical flip, the texture path of
// Texture2d::new (src/texture.rs:487-496: image::open, flipv, to_rgba8) used
// for the AO blue-noise texture.  8- and 16-bit grey / grey+alpha / RGB /
// RGBA, non-interlaced; 16-bit channels become (v + 128) / 257.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sol {

struct ImageRgba8 {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> rgba;  // row-major, 4 bytes per texel
};

// throws std::runtime_error on unsupported or malformed files
ImageRgba8 load_png_rgba8_flipv(const std::string &path);

}  // namespace sol
