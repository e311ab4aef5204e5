// Planar fp32 image, as sol::Image (reference include/sol/image.h:18-98, src/image.cpp:13-41).
// Pixel (0, 0) is the top-left corner. Files: OpenEXR (sol's default format, written as sol's
// exr::save writes it: uncompressed 32-bit float scanlines, channels B, G, R) and PFM; PNG, JPEG
// and TIFF need codec libraries this library does not link, so saving in them fails, as sol's does
// when it is built without them.
#ifndef SOL_IMAGE_H
#define SOL_IMAGE_H

#include <cassert>
#include <cstddef>
#include <memory>
#include <string_view>
#include <vector>

#include "sol/color.h"

namespace sol {

struct Image {
    using Channel = std::unique_ptr<float[]>;

    Image() = default;
    Image(const Image&) = delete;
    Image(Image&&) = default;
    Image& operator=(Image&&) = default;
    Image(size_t width, size_t height, size_t channel_count);

    size_t width() const { return width_; }
    size_t height() const { return height_; }
    size_t channel_count() const { return channels_.size(); }

    RgbColor rgb_at(size_t x, size_t y) const {
        assert(channel_count() == 3);
        size_t i = y * width_ + x;
        return RgbColor(channels_[0][i], channels_[1][i], channels_[2][i]);
    }
    void accumulate(size_t x, size_t y, const RgbColor& c) {
        assert(channel_count() == 3);
        size_t i = y * width_ + x;
        channels_[0][i] += c.r;
        channels_[1][i] += c.g;
        channels_[2][i] += c.b;
    }

    Channel& channel(size_t i) { return channels_[i]; }
    const Channel& channel(size_t i) const { return channels_[i]; }

    enum class Format { Auto, Png, Jpeg, Tiff, Exr, Pfm };

    void clear(float value = 0.0f);
    void scale(float value);
    /// Image::save (image.cpp:26-37): Auto means EXR. False if the file cannot be written or the
    /// format has no writer here (PNG, JPEG, TIFF).
    bool save(std::string_view path, Format format = Format::Auto) const;
    // OpenEXR, single part, scanline, uncompressed, FLOAT channels (exr.cpp:48-100).
    bool save_exr(std::string_view path) const;
    // Portable float map (RGB, bottom-to-top rows as the format requires).
    bool save_pfm(std::string_view path) const;

private:
    size_t width_ = 0, height_ = 0;
    std::vector<Channel> channels_;
};

} // namespace sol

#endif
