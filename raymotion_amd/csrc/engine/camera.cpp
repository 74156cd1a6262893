// Camera focal length and rotation matrix (src/engine/camera.cpp:3-22 of the
// reference; same formulas, so matrices match the reference bit for bit).
#include "raymotion/pathtrace.hpp"

namespace raymotion {

void Camera::update_focal() { focal = 240.0 / tan_deg(fov * 0.5); }

void Camera::update_trig() {
    const double cx = cos_deg(dirx), sx = sin_deg(dirx);
    const double cy = cos_deg(diry), sy = sin_deg(diry);
    const double cz = cos_deg(dirz), sz = sin_deg(dirz);
    m0 = cy * cz + sy * (sx * sz);
    m1 = -cy * sz + sy * (sx * cz);
    m2 = sy * cx;
    m3 = cx * sz;
    m4 = cx * cz;
    m5 = -sx;
    m6 = -sy * cz + cy * (sx * sz);
    m7 = sy * sz + cy * (sx * cz);
    m8 = cy * cx;
}

}  // namespace raymotion
