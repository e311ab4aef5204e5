// Image storage (planar channels: construction, clear, scale; reference src/image.cpp:13-23) and
// image files: PFM, and the OpenEXR writer and format dispatch (reference src/image.cpp:26-37,
// src/formats/exr.cpp:48-100). The EXR file follows the OpenEXR file layout: magic number, version
// 2, a header of attributes, the scanline offset table, then one block per scanline.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sol/image.h"

namespace sol {

namespace {

struct Bytes {
    std::vector<unsigned char> b;
    void raw(const void* p, size_t n) {
        auto* c = static_cast<const unsigned char*>(p);
        b.insert(b.end(), c, c + n);
    }
    void u8(uint8_t v) { b.push_back(v); }
    void i32(int32_t v) { for (int i = 0; i < 4; ++i) b.push_back(static_cast<unsigned char>(uint32_t(v) >> (8 * i))); }
    void u64(uint64_t v) { for (int i = 0; i < 8; ++i) b.push_back(static_cast<unsigned char>(v >> (8 * i))); }
    void f32(float v) {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        i32(static_cast<int32_t>(u));
    }
    void str(const std::string& s) { raw(s.c_str(), s.size() + 1); }
    // name, type, size, value
    void attr(const std::string& name, const std::string& type, const Bytes& value) {
        str(name);
        str(type);
        i32(static_cast<int32_t>(value.b.size()));
        raw(value.b.data(), value.b.size());
    }
};

} // namespace

Image::Image(size_t width, size_t height, size_t channel_count) : width_(width), height_(height) {
    channels_.resize(channel_count);
    for (auto& c : channels_) {
        c = std::make_unique<float[]>(width * height);
        std::fill(c.get(), c.get() + width * height, 0.0f);
    }
}

void Image::clear(float value) {
    for (auto& c : channels_) std::fill(c.get(), c.get() + width_ * height_, value);
}

void Image::scale(float value) {
    for (auto& c : channels_)
        for (size_t i = 0; i < width_ * height_; ++i) c[i] *= value;
}

bool Image::save_pfm(std::string_view path) const {
    if (channel_count() != 3) return false;
    std::string p(path);
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) return false;
    std::fprintf(f, "PF\n%zu %zu\n-1.0\n", width_, height_);
    std::vector<float> row(width_ * 3);
    for (size_t y = height_; y-- > 0;) {
        for (size_t x = 0; x < width_; ++x)
            for (int c = 0; c < 3; ++c) row[x * 3 + c] = channels_[c][y * width_ + x];
        std::fwrite(row.data(), sizeof(float), row.size(), f);
    }
    return std::fclose(f) == 0;
}

bool Image::save_exr(std::string_view path) const {
    const size_t n = channel_count();
    if (n == 0 || width_ == 0 || height_ == 0 || width_ > 0x7FFFFFFF || height_ > 0x7FFFFFFF) return false;
    // Channel names as sol gives them (exr.cpp:73-84): B, G, R for the first three channels of an
    // RGB(A) image, A for a fourth, Channel_i otherwise; the file lists them by name.
    std::vector<std::pair<std::string, size_t>> chans;
    for (size_t i = 0; i < n; ++i) {
        std::string name = "Channel_" + std::to_string(i);
        if (n >= 3 && i < 3) name = i == 0 ? "R" : (i == 1 ? "G" : "B");
        if (n >= 4 && i == 3) name = "A";
        chans.emplace_back(name, i);
    }
    std::sort(chans.begin(), chans.end());

    Bytes f;
    f.i32(20000630);   // magic
    f.i32(2);          // version 2, single-part scanline, short names
    Bytes chlist;
    for (auto& [name, i] : chans) {
        chlist.str(name);
        chlist.i32(2);                 // pixel type FLOAT
        chlist.u8(0);                  // pLinear
        chlist.u8(0); chlist.u8(0); chlist.u8(0);
        chlist.i32(1);                 // x sampling
        chlist.i32(1);                 // y sampling
    }
    chlist.u8(0);
    f.attr("channels", "chlist", chlist);
    Bytes comp;
    comp.u8(0);                        // NO_COMPRESSION
    f.attr("compression", "compression", comp);
    Bytes window;
    window.i32(0); window.i32(0);
    window.i32(static_cast<int32_t>(width_ - 1)); window.i32(static_cast<int32_t>(height_ - 1));
    f.attr("dataWindow", "box2i", window);
    f.attr("displayWindow", "box2i", window);
    Bytes order;
    order.u8(0);                       // INCREASING_Y
    f.attr("lineOrder", "lineOrder", order);
    Bytes one;
    one.f32(1.0f);
    f.attr("pixelAspectRatio", "float", one);
    Bytes center;
    center.f32(0.0f); center.f32(0.0f);
    f.attr("screenWindowCenter", "v2f", center);
    f.attr("screenWindowWidth", "float", one);
    f.u8(0);                           // end of header

    const uint64_t line_bytes = uint64_t(n) * width_ * sizeof(float);
    const uint64_t table_end = f.b.size() + uint64_t(height_) * 8;
    for (size_t y = 0; y < height_; ++y) f.u64(table_end + uint64_t(y) * (8 + line_bytes));

    std::string p(path);
    FILE* out = std::fopen(p.c_str(), "wb");
    if (!out) return false;
    bool ok = std::fwrite(f.b.data(), 1, f.b.size(), out) == f.b.size();
    std::vector<unsigned char> block(8 + line_bytes);
    for (size_t y = 0; ok && y < height_; ++y) {
        int32_t yy = static_cast<int32_t>(y), size = static_cast<int32_t>(line_bytes);
        std::memcpy(block.data(), &yy, 4);
        std::memcpy(block.data() + 4, &size, 4);
        size_t off = 8;
        for (auto& [name, i] : chans) {
            std::memcpy(block.data() + off, channels_[i].get() + y * width_, width_ * sizeof(float));
            off += width_ * sizeof(float);
        }
        ok = std::fwrite(block.data(), 1, block.size(), out) == block.size();
    }
    return std::fclose(out) == 0 && ok;
}

bool Image::save(std::string_view path, Format format) const {
    switch (format) {
        case Format::Auto:
        case Format::Exr: return save_exr(path);
        case Format::Pfm: return save_pfm(path);
        default: return false;    // PNG, JPEG, TIFF: no codec library is linked
    }
}

} // namespace sol
