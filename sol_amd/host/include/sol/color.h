// RGB colour, as sol::RgbColor (reference include/sol/color.h:11-64). The reference's
// operator/=(float) multiplies instead of dividing (color.h:23); this one divides.
#ifndef SOL_COLOR_H
#define SOL_COLOR_H

namespace sol {

struct RgbColor {
    float r = 0, g = 0, b = 0;

    RgbColor() = default;
    explicit constexpr RgbColor(float rgb) : r(rgb), g(rgb), b(rgb) {}
    constexpr RgbColor(float r, float g, float b) : r(r), g(g), b(b) {}

    RgbColor operator+(const RgbColor& o) const { return RgbColor(r + o.r, g + o.g, b + o.b); }
    RgbColor operator-(const RgbColor& o) const { return RgbColor(r - o.r, g - o.g, b - o.b); }
    RgbColor operator*(const RgbColor& o) const { return RgbColor(r * o.r, g * o.g, b * o.b); }
    RgbColor operator*(float s) const { return RgbColor(r * s, g * s, b * s); }
    RgbColor operator/(float s) const { return *this * (1.0f / s); }
    RgbColor& operator+=(const RgbColor& o) { return *this = *this + o; }
    RgbColor& operator*=(const RgbColor& o) { return *this = *this * o; }
    RgbColor& operator*=(float s) { return *this = *this * s; }
    RgbColor& operator/=(float s) { return *this = *this / s; }
    bool operator==(const RgbColor& o) const { return r == o.r && g == o.g && b == o.b; }

    float luminance() const { return r * 0.2126f + g * 0.7152f + b * 0.0722f; }
    bool is_black() const { return r == 0.0f && g == 0.0f && b == 0.0f; }
    static constexpr RgbColor black() { return RgbColor(0.0f); }
    static constexpr RgbColor constant(float c) { return RgbColor(c); }
};

using Color = RgbColor;

} // namespace sol

#endif
