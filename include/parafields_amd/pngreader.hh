// pngreader.hh -- pixel values of a PNG file for the image trend component
// (the reference's PNGReader, pngreader.hh:1-113). Reading needs libpng: a
// program that defines PARAFIELDS_AMD_HAVE_PNG (the reference's HAVE_PNG)
// includes <png.h> and links libpng; without it the constructor throws, as
// the reference's does without HAVE_PNG.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef PARAFIELDS_AMD_HAVE_PNG
#include <png.h>
#endif

namespace parafields_amd {

class PNGReader
{
  unsigned int offset_ = 0, width_ = 0, height_ = 0, channels_ = 0;
  std::vector<std::vector<unsigned char>> rows_;

public:
  // the rows of the image as libpng hands them out after interlace
  // handling, one byte per sample of 8-bit images (pngreader.hh:27-74)
  explicit PNGReader(const std::string& filename, unsigned int channelOffset = 0)
    : offset_(channelOffset)
  {
#ifdef PARAFIELDS_AMD_HAVE_PNG
    FILE* file = std::fopen(filename.c_str(), "rb");
    if (!file)
      throw std::runtime_error("opening PNG file failed");
    unsigned char header[8];
    if (std::fread(header, 1, 8, file) != 8 || png_sig_cmp(header, 0, 8)) {
      std::fclose(file);
      throw std::runtime_error("file not recognized as PNG");
    }
    png_structp png = png_create_read_struct(PNG_LIBPNG_VER_STRING, nullptr, nullptr, nullptr);
    png_infop info = png ? png_create_info_struct(png) : nullptr;
    if (!png || !info) {
      png_destroy_read_struct(&png, nullptr, nullptr);
      std::fclose(file);
      throw std::runtime_error("couldn't create the PNG reader");
    }
    std::vector<png_bytep> ptrs; // declared before the error return point
    if (setjmp(png_jmpbuf(png))) {
      png_destroy_read_struct(&png, &info, nullptr);
      std::fclose(file);
      throw std::runtime_error("reading PNG file failed");
    }
    png_init_io(png, file);
    png_set_sig_bytes(png, 8);
    png_read_info(png, info);
    width_ = png_get_image_width(png, info);
    height_ = png_get_image_height(png, info);
    channels_ = png_get_channels(png, info);
    if (offset_ >= channels_) {
      png_destroy_read_struct(&png, &info, nullptr);
      std::fclose(file);
      throw std::runtime_error("chosen channel offset exceeds number of channels in PNG");
    }
    (void)png_set_interlace_handling(png);
    png_read_update_info(png, info);
    const size_t rowBytes = png_get_rowbytes(png, info);
    rows_.assign(height_, std::vector<unsigned char>(rowBytes));
    ptrs.resize(height_);
    for (unsigned int r = 0; r < height_; r++)
      ptrs[r] = rows_[r].data();
    png_read_image(png, ptrs.data());
    png_destroy_read_struct(&png, &info, nullptr);
    std::fclose(file);
#else
    (void)filename;
    throw std::runtime_error("missing support for reading PNG files");
#endif
  }

  unsigned int width() const { return width_; }
  unsigned int height() const { return height_; }
  unsigned int channels() const { return channels_; }

  // 1 - value / 255 of the pixel holding `coord` in a domain of size
  // `extent`, the image covering the domain with its first row at the top
  // (pngreader.hh:90-103)
  template <typename C1, typename C2>
  double evaluate(const C1& coord, const C2& extent) const
  {
    const unsigned int row = height_ - 1 - int(coord[1] / double(extent[1]) * height_);
    const unsigned int col = int(coord[0] / double(extent[0]) * width_);
    return 1. - rows_.at(row).at(col * channels_ + offset_) / 255.;
  }
};

} // namespace parafields_amd
