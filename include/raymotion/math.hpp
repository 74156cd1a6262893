#pragma once
// Scalar helpers and the double-precision vector used by the scene API.
// Interface follows include/raymotion/math.hpp of the reference (Vec3, degree
// trigonometry, gobo_mod, gobo_normalize, Rng); the code is this project's own.
#include <cmath>
#include <cstdint>

namespace raymotion {

inline constexpr double kPi = 3.141592653589793;
inline constexpr double kFarClip = 1e30;
inline constexpr double kEps = 1e-8;
inline constexpr double kDenomEps = 1e-12;

inline double deg2rad(double d) { return d * kPi / 180.0; }
inline double rad2deg(double r) { return r * 180.0 / kPi; }

// Angles in the scene language are degrees.
inline double cos_deg(double d) { return std::cos(deg2rad(d)); }
inline double sin_deg(double d) { return std::sin(deg2rad(d)); }
inline double tan_deg(double d) { return std::tan(deg2rad(d)); }
inline double atan_deg(double v) { return rad2deg(std::atan(v)); }

// Floored modulo: for b > 0 the result lies in [0, b).
inline double gobo_mod(double a, double b) {
    if (b == 0) return 0;
    return a - std::floor(a / b) * b;
}

struct Vec3 {
    double x = 0, y = 0, z = 0;
    Vec3() = default;
    Vec3(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
    Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vec3 operator*(double s) const { return {x * s, y * s, z * s}; }
    Vec3 operator/(double s) const { return {x / s, y / s, z / s}; }
    Vec3& operator+=(const Vec3& o) { x += o.x; y += o.y; z += o.z; return *this; }
};

inline double dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(const Vec3& a, const Vec3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double length(const Vec3& v) { return std::sqrt(dot(v, v)); }
inline Vec3 normalize(const Vec3& v) {
    const double l = length(v);
    return l < 1e-30 ? Vec3{} : v / l;
}
// x*d, y*d, z*d with d = 1/sqrt(x^2 + (y^2 + z^2)) (note the grouping).
inline Vec3 gobo_normalize(double x, double y, double z) {
    const double d = 1.0 / std::sqrt(x * x + (y * y + z * z));
    return {x * d, y * d, z * d};
}

// Host-side uniform generator for scene programs (xorshift64* seeded through
// splitmix64). Render sampling uses its own per-pixel PCG streams instead.
class Rng {
public:
    explicit Rng(uint64_t s = 12345) { seed(s); }
    void seed(uint64_t s) {
        uint64_t z = s + 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        state_ = (z ^ (z >> 31)) | 1ULL;
    }
    double uniform01() {
        state_ ^= state_ >> 12;
        state_ ^= state_ << 25;
        state_ ^= state_ >> 27;
        return double((state_ * 0x2545F4914F6CDD1DULL) >> 11) * (1.0 / 9007199254740992.0);
    }
    double uniform(double lo, double hi) {
        if (lo > hi) { double t = lo; lo = hi; hi = t; }
        return lo + (hi - lo) * uniform01();
    }
private:
    uint64_t state_ = 1;
};

}  // namespace raymotion
