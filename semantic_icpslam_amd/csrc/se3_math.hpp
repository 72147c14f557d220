// se3_math.hpp — FP64 6x6 solve and SE(3) algebra of one Gauss-Newton
// iteration (Tracking/Tracking.cc:493-498, 505-511), compiled for both the
// host and the device.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define SE3_HD __host__ __device__ inline
#else
#define SE3_HD inline
#endif

namespace sicp {

// CreateMatrix<6,7> (Tracking/Reduction.cu:6-16): the 27 upper-triangle sums
// become the mirrored A (row-major) and b.
SE3_HD void create_matrix(const float* s, double* A, double* b) {
    int shift = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j) {
            double value = (double)s[shift++];
            if (j == 6)
                b[i] = value;
            else
                A[j * 6 + i] = A[i * 6 + j] = value;
        }
}

// x = A^-1 b by LDL^T without pivoting (a zero pivot gives a zero solution
// component, as Eigen's LDLT::solve, Tracking.cc:493). Returns false on a
// zero pivot. One reciprocal per pivot: the solve runs on one GPU thread per
// iteration, where every FP64 division is a long dependent sequence.
SE3_HD bool ldlt_solve6(const double* A, const double* b, double* x) {
    double L[36], D[6], Dinv[6];
    for (int i = 0; i < 36; ++i) L[i] = 0.0;
    for (int j = 0; j < 6; ++j) {
        double d = A[j * 6 + j];
        for (int k = 0; k < j; ++k) d -= L[j * 6 + k] * L[j * 6 + k] * D[k];
        D[j] = d;
        Dinv[j] = (d != 0.0) ? 1.0 / d : 0.0;
        L[j * 6 + j] = 1.0;
        for (int i = j + 1; i < 6; ++i) {
            double s = A[i * 6 + j];
            for (int k = 0; k < j; ++k) s -= L[i * 6 + k] * L[j * 6 + k] * D[k];
            L[i * 6 + j] = s * Dinv[j];
        }
    }
    double z[6];
    for (int i = 0; i < 6; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[i * 6 + k] * z[k];
        z[i] = s;
    }
    bool ok = true;
    for (int i = 0; i < 6; ++i) {
        z[i] = z[i] * Dinv[i];
        if (D[i] == 0.0) ok = false;
    }
    for (int i = 5; i >= 0; --i) {
        double s = z[i];
        for (int k = i + 1; k < 6; ++k) s -= L[k * 6 + i] * x[k];
        x[i] = s;
    }
    return ok;
}

// exp of the twist xi = (v, w): rotation from the unit quaternion
// (cos(t/2), sin(t/2) w / t), translation V v with
// V = I + (1 - cos t)/t^2 W + (t - sin t)/t^3 W^2 (small-angle series below
// 1e-10) — the closed form Sophus::SE3d::exp evaluates (Tracking.cc:494).
// One sincos of t/2 serves all four trigonometric values:
// 1 - cos t = 2 sin^2(t/2) and sin t = 2 sin(t/2) cos(t/2).
SE3_HD void se3_exp(const double* xi, double* T) {
    const double* v = xi;
    const double* w = xi + 3;
    const double eps = 1e-10;
    double theta_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    double theta = sqrt(theta_sq);
    double half = 0.5 * theta;
    double imag, real, sh = 0.0, ch = 1.0, inv_t = 0.0;
    if (theta_sq < eps * eps) {
        double t4 = theta_sq * theta_sq;
        imag = 0.5 - theta_sq / 48.0 + t4 / 3840.0;
        real = 1.0 - theta_sq / 8.0 + t4 / 384.0;
    } else {
        sincos(half, &sh, &ch);
        inv_t = 1.0 / theta;
        imag = sh * inv_t;
        real = ch;
    }
    double qw = real, qx = imag * w[0], qy = imag * w[1], qz = imag * w[2];
    double tx = 2 * qx, ty = 2 * qy, tz = 2 * qz;
    double twx = tx * qw, twy = ty * qw, twz = tz * qw;
    double txx = tx * qx, txy = ty * qx, txz = tz * qx;
    double tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
    double Rm[9] = {1 - (tyy + tzz), txy - twz,       txz + twy,
                    txy + twz,       1 - (txx + tzz), tyz - twx,
                    txz - twy,       tyz + twx,       1 - (txx + tyy)};
    double V[9];
    if (theta < eps) {
        for (int i = 0; i < 9; ++i) V[i] = Rm[i];
    } else {
        double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
        double W2[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                W2[i * 3 + j] = W[i * 3 + 0] * W[0 * 3 + j] + W[i * 3 + 1] * W[1 * 3 + j] +
                                W[i * 3 + 2] * W[2 * 3 + j];
        if (inv_t == 0.0) {  // theta^2 rounded below eps^2 but theta did not
            sincos(half, &sh, &ch);
            inv_t = 1.0 / theta;
        }
        const double st = sh * inv_t;
        double c1 = 2.0 * st * st;                                    // (1 - cos t) / t^2
        double c2 = (theta - 2.0 * sh * ch) * (inv_t * inv_t * inv_t);  // (t - sin t) / t^3
        for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0 ? 1.0 : 0.0) + c1 * W[i] + c2 * W2[i];
    }
    for (int i = 0; i < 3; ++i) {
        T[i * 4 + 0] = Rm[i * 3 + 0];
        T[i * 4 + 1] = Rm[i * 3 + 1];
        T[i * 4 + 2] = Rm[i * 3 + 2];
        T[i * 4 + 3] = V[i * 3 + 0] * v[0] + V[i * 3 + 1] * v[1] + V[i * 3 + 2] * v[2];
    }
    T[12] = T[13] = T[14] = 0.0;
    T[15] = 1.0;
}

SE3_HD void mat4_mul(const double* A, const double* B, double* C) {
    double t[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            t[i * 4 + j] = A[i * 4 + 0] * B[0 * 4 + j] + A[i * 4 + 1] * B[1 * 4 + j] +
                           A[i * 4 + 2] * B[2 * 4 + j] + A[i * 4 + 3] * B[3 * 4 + j];
    for (int i = 0; i < 16; ++i) C[i] = t[i];
}

SE3_HD void rigid_inverse(const double* T, double* Ti) {
    double t[16];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[i * 4 + j] = T[j * 4 + i];
    for (int i = 0; i < 3; ++i)
        t[i * 4 + 3] = -(t[i * 4 + 0] * T[3] + t[i * 4 + 1] * T[7] + t[i * 4 + 2] * T[11]);
    t[12] = t[13] = t[14] = 0.0;
    t[15] = 1.0;
    for (int i = 0; i < 16; ++i) Ti[i] = t[i];
}

SE3_HD void set_identity(double* T) {
    for (int i = 0; i < 16; ++i) T[i] = (i % 5 == 0) ? 1.0 : 0.0;
}

// Norm of the sines of the XYZ Euler angles of T's rotation, and norm of
// T's translation (Tracking.cc:505-511); both Euler solutions give the same
// sine norm.
SE3_HD void euler_sin_norm(const double* T, double* rot_norm, double* trans_norm) {
    double a = atan2(-T[6], T[10]);
    double b = atan2(T[2], sqrt(T[0] * T[0] + T[1] * T[1]));
    double c = atan2(-T[1], T[0]);
    double sa = sin(a), sb = sin(b), sc = sin(c);
    *rot_norm = sqrt(sa * sa + sb * sb + sc * sc);
    *trans_norm = sqrt(T[3] * T[3] + T[7] * T[7] + T[11] * T[11]);
}

// Float view of a double pose, as Frame::GpuRotation / GpuInvRotation /
// GpuTranslation (Core/Frame.cc:626-645).
template <class P> SE3_HD void pose_to_float(const double* T, P* out) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            out->R[i * 3 + j] = (float)T[i * 4 + j];
            out->Rinv[i * 3 + j] = (float)T[j * 4 + i];
        }
    out->t[0] = (float)T[3];
    out->t[1] = (float)T[7];
    out->t[2] = (float)T[11];
}

}  // namespace sicp
