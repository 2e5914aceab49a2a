// Small-vector/tensor arithmetic used by the device kernels.  Operation
// order follows OpenFOAM's Vector/Tensor operators (x+y then +z for inner
// products, row-major tensors xx xy xz yx yy yz zx zy zz), so that face
// fluxes and Jacobian blocks come out bit-identical to a sequential CPU
// evaluation of the same expressions when both are built without
// floating-point contraction.
#pragma once
#include <hip/hip_runtime.h>

namespace icsf {

#define HD __host__ __device__ __forceinline__

constexpr double kSMALL = 1.0e-15;
constexpr double kVSMALL = 1.0e-300;
constexpr double kROOTVSMALL = 1.0e-150;
constexpr double kGREAT = 1.0e15;

struct V3 {
    double x, y, z;
};
HD V3 mk3(double x, double y, double z) { return V3{x, y, z}; }
HD V3 ld3(const double* p) { return V3{p[0], p[1], p[2]}; }
HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
HD V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
HD V3 operator*(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
HD V3 operator/(V3 a, double s) { return V3{a.x / s, a.y / s, a.z / s}; }
HD double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HD double magSqr(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
HD double mag(V3 a) { return sqrt(magSqr(a)); }
HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
HD double sqr(double x) { return x * x; }
// transform(R, v) = R & v for a row-major rotation tensor R
HD V3 rot3(const double* R, V3 v) {
    return V3{R[0] * v.x + R[1] * v.y + R[2] * v.z, R[3] * v.x + R[4] * v.y + R[5] * v.z,
              R[6] * v.x + R[7] * v.y + R[8] * v.z};
}

struct T9 {
    double t[9];
};
HD T9 outer(V3 a, V3 b) {
    return T9{{a.x * b.x, a.x * b.y, a.x * b.z, a.y * b.x, a.y * b.y, a.y * b.z, a.z * b.x, a.z * b.y, a.z * b.z}};
}
HD T9 operator+(const T9& a, const T9& b) {
    T9 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.t[i] = a.t[i] + b.t[i];
    return r;
}
HD T9 operator-(const T9& a, const T9& b) {
    T9 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.t[i] = a.t[i] - b.t[i];
    return r;
}
HD T9 operator*(double s, const T9& a) {
    T9 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.t[i] = s * a.t[i];
    return r;
}
HD T9 operator*(const T9& a, double s) {
    T9 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.t[i] = a.t[i] * s;
    return r;
}
HD V3 tdotv(const T9& T, V3 v) {
    return V3{T.t[0] * v.x + T.t[1] * v.y + T.t[2] * v.z, T.t[3] * v.x + T.t[4] * v.y + T.t[5] * v.z,
              T.t[6] * v.x + T.t[7] * v.y + T.t[8] * v.z};
}
HD V3 vdott(V3 v, const T9& T) {
    return V3{v.x * T.t[0] + v.y * T.t[3] + v.z * T.t[6], v.x * T.t[1] + v.y * T.t[4] + v.z * T.t[7],
              v.x * T.t[2] + v.y * T.t[5] + v.z * T.t[8]};
}
HD T9 tdott(const T9& A, const T9& B) {
    T9 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.t[3 * i + j] = A.t[3 * i] * B.t[j] + A.t[3 * i + 1] * B.t[3 + j] + A.t[3 * i + 2] * B.t[6 + j];
    return r;
}
HD T9 transposeT(const T9& A) { return T9{{A.t[0], A.t[3], A.t[6], A.t[1], A.t[4], A.t[7], A.t[2], A.t[5], A.t[8]}}; }
HD T9 dev2(const T9& A) {
    T9 r = A;
    const double tr = A.t[0] + A.t[4] + A.t[8];
    const double s = (2.0 / 3.0) * tr;
    r.t[0] = A.t[0] - s;
    r.t[4] = A.t[4] - s;
    r.t[8] = A.t[8] - s;
    return r;
}
// (I - 2 n n) & v : the basicSymmetry reflection
HD V3 reflect(V3 n, V3 v) {
    T9 Tr{{1 - 2.0 * n.x * n.x, -2.0 * n.x * n.y, -2.0 * n.x * n.z, -2.0 * n.y * n.x, 1 - 2.0 * n.y * n.y,
           -2.0 * n.y * n.z, -2.0 * n.z * n.x, -2.0 * n.z * n.y, 1 - 2.0 * n.z * n.z}};
    return tdotv(Tr, v);
}

#undef HD
}  // namespace icsf
