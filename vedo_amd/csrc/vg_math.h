// vg_math.h -- FP64 vector arithmetic for the DEM kernels.
//
// Every helper rounds in the same order as the reference's aries::Vector3d
// (Vector3d.h:15-316), and the library is compiled with -ffp-contract=off,
// so a product a*b+c is two roundings as in the reference's SSE2 build.
// That makes the GPU contact geometry and spring-dashpot forces reproduce the
// reference's FP64 bits for the same inputs and summation order.
#pragma once

#include <hip/hip_runtime.h>

namespace vg {

struct d3 {
    double x, y, z;
};

__host__ __device__ inline d3 mk(double x, double y, double z) { return d3{x, y, z}; }
__host__ __device__ inline d3 operator+(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline d3 operator-(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline d3 operator-(d3 a) { return mk(-a.x, -a.y, -a.z); }
// scalar * vector: (s*v.x, s*v.y, s*v.z)  (Vector3d.h:313-316)
__host__ __device__ inline d3 smul(double s, d3 v) { return mk(s * v.x, s * v.y, s * v.z); }
// vector * scalar: (v.x*s, ...)  (Vector3d.h:268-271)
__host__ __device__ inline d3 muls(d3 v, double s) { return mk(v.x * s, v.y * s, v.z * s); }
__host__ __device__ inline double dot(d3 a, d3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
__host__ __device__ inline d3 cross(d3 a, d3 b)
{
    return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
__host__ __device__ inline double len(d3 a) { return sqrt((a.x * a.x) + (a.y * a.y) + (a.z * a.z)); }
// direction(): zero for the zero vector
__host__ __device__ inline d3 dir(d3 a)
{
    double l = len(a);
    if (l != 0) return mk(a.x / l, a.y / l, a.z / l);
    return mk(0.0, 0.0, 0.0);
}
__host__ __device__ inline d3 project_on(d3 a, d3 v) { return smul(dot(a, v) / len(v), dir(v)); }

// rotate_around (Vector3d.h:152-163): quaternion-style rotation of p by the
// rotation vector va.
__device__ inline d3 rotate_around(d3 p, d3 va)
{
    d3 u = dir(va);
    double a = len(va);
    double s = cos(0.5 * a);
    double t = sin(0.5 * a);
    d3 v = muls(u, t);
    d3 t1 = smul(s * s, p);
    d3 t2 = smul(dot(p, v), v);
    d3 t3 = smul(2.0 * s, cross(v, p));
    d3 t4 = cross(v, cross(v, p));
    return ((t1 + t2) + t3) + t4;
}

// rotate_around applied to two vectors with the same rotation vector: the
// shared terms (direction, half-angle cos/sin) are identical in both calls
// of the reference (DOSphere.cpp:41), so computing them once is exact.
__device__ inline void rotate_pair(d3 &p1, d3 &p2, d3 va)
{
    d3 u = dir(va);
    double a = len(va);
    double s = cos(0.5 * a);
    double t = sin(0.5 * a);
    d3 v = muls(u, t);
    const double ss = s * s, s2 = 2.0 * s;
    d3 t1 = smul(ss, p1), t2 = smul(dot(p1, v), v), t3 = smul(s2, cross(v, p1)), t4 = cross(v, cross(v, p1));
    p1 = ((t1 + t2) + t3) + t4;
    t1 = smul(ss, p2); t2 = smul(dot(p2, v), v); t3 = smul(s2, cross(v, p2)); t4 = cross(v, cross(v, p2));
    p2 = ((t1 + t2) + t3) + t4;
}

__device__ inline d3 ld3(const double4 &v) { return mk(v.x, v.y, v.z); }

}  // namespace vg
