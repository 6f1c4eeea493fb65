// analytical_flow.hpp -- Taylor-Green vortex (include/analytical_flow.hpp:12-31)
#pragma once

#include <cmath>

#include "device_vector_alias.hpp"

constexpr float TWO_PI = 6.28318530717958647692f;

// u = sin(2 pi x) cos(2 pi y), v = -cos(2 pi x) sin(2 pi y)
HD inline float2 taylor_green_velocity(float x, float y)
{
    return make_float2(sinf(TWO_PI * x) * cosf(TWO_PI * y), -cosf(TWO_PI * x) * sinf(TWO_PI * y));
}
HD inline float2 taylor_green_velocity(const float2 &p) { return taylor_green_velocity(p.x, p.y); }

// omega = 4 pi sin(2 pi x) sin(2 pi y)
HD inline float taylor_green_vorticity(float x, float y)
{
    return 4.0f * 3.14159265358979323846f * sinf(TWO_PI * x) * sinf(TWO_PI * y);
}
HD inline float taylor_green_vorticity(const float2 &p) { return taylor_green_vorticity(p.x, p.y); }
