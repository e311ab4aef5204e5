// Planar fp32 image, as sol::Image (reference include/sol/image.h:18-98, src/image.cpp:13-23).
// Pixel (0, 0) is the top-left corner. Only the in-memory operations are on the path; a PFM
// writer stands in for the EXR/PNG codecs, whose libraries are absent.
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

    void clear(float value = 0.0f);
    void scale(float value);
    // Portable float map (RGB, bottom-to-top rows as the format requires).
    bool save_pfm(std::string_view path) const;

private:
    size_t width_ = 0, height_ = 0;
    std::vector<Channel> channels_;
};

} // namespace sol

#endif
