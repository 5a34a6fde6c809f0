// Projection image decoding for the CtVolume facade.
//
// Ct decodes every projection with cv::imread(GRAYSCALE | ANYDEPTH) and accepts the
// depths 8U, 16U and 32F (Ct/CtVolume.cpp:20-33, 247).  OpenCV is not part of this
// build, so the formats CT projection stacks are stored in are decoded here:
//   - PGM (P5 binary, P2 ASCII), 8- or 16-bit;
//   - TIFF (baseline, strips, little or big endian), 8/16-bit unsigned or 32-bit float,
//     uncompressed, PackBits, LZW or Deflate, with or without horizontal predictor;
//   - PNG, 8- or 16-bit, grayscale (with or without alpha) or RGB(A) converted to gray
//     with the BT.601 weights, non-interlaced.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ct {

enum class PixelDepth { U8 = 0, U16 = 1, F32 = 2 };

struct Image {
    int width = 0;
    int height = 0;
    PixelDepth depth = PixelDepth::U16;
    std::vector<uint8_t> data;  // row-major, native byte order

    size_t bytesPerPixel() const { return depth == PixelDepth::U8 ? 1 : 4 - 2 * (depth == PixelDepth::U16); }
    bool empty() const { return width == 0 || height == 0; }
};

// Decodes the file at `path`.  Returns false and sets `err` on failure.
bool readImage(const std::string& path, Image& out, std::string& err);

// Writes a 16-bit binary PGM (big-endian samples, maxval 65535).
bool writePgm16(const std::string& path, const uint16_t* pixels, int width, int height, std::string& err);

}  // namespace ct
