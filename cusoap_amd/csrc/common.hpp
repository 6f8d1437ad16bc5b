// Shared device code of the cusoap_amd kernels: error plumbing, real solid
// harmonics, radial functions and the weighting function.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/cusoap_amd.h"
#include "sh_constants.inc"

namespace cusoap {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define CUSOAP_HIP(expr)                                                              \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess)                                                         \
            return ::cusoap::fail(CUSOAP_ERR_HIP, std::string(#expr) + ": " +         \
                                                      hipGetErrorString(_e));         \
    } while (0)

#define CUSOAP_LAUNCH_CHECK()                                                         \
    do {                                                                              \
        hipError_t _e = hipGetLastError();                                            \
        if (_e != hipSuccess)                                                         \
            return ::cusoap::fail(CUSOAP_ERR_HIP, std::string("kernel launch: ") +    \
                                                      hipGetErrorString(_e));         \
    } while (0)

// Optional per-kernel timing with HIP events recorded on the launch stream.
void prof_begin(const char* name, hipStream_t st);
void prof_end(const char* name, hipStream_t st);

constexpr int kMaxL = 9;        // compiled l_max range 0..kMaxL
constexpr int kMaxN = 16;       // n_max range 1..kMaxN
constexpr int kMaxS = 64;       // species

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Extended index ext = img * n_atoms + atom (0 <= ext < 2^31) split without an
// integer division.  inv_n = 1.0 / n_atoms in float64: ext * inv_n has a relative
// error below 2^-51, so for ext < 2^31 it never reaches the next integer above
// ext / n_atoms, and it truncates to the true quotient except at exact multiples,
// where it can be one less; the remainder check corrects that case.
__device__ __forceinline__ void split_ext(int ext, int n_atoms, double inv_n, int& img, int& atom) {
    int q = (int)((double)ext * inv_n);
    int r = ext - q * n_atoms;
    if (r >= n_atoms) {
        ++q;
        r -= n_atoms;
    }
    img = q;
    atom = r;
}

// ---------------------------------------------------------------------------
// real solid harmonics r^l Y_lm(r) and Cartesian gradients
// ---------------------------------------------------------------------------
// Associated-Legendre recurrence in (z, r^2) for Q_l^m, with (x + i y)^m for the
// azimuthal part:  Y_l0 = K_l0 Q_l^0,  Y_l,+m = K_lm Q_l^m Re(x+iy)^m,
// Y_l,-m = K_lm Q_l^m Im(x+iy)^m.  Gradients by forward differentiation of the
// recurrence (dQ/dz at fixed r^2 and dQ/dr^2).  `put(lm, v, gx, gy, gz)` is
// called once per lm.
template <int LMAX, bool GRAD, typename T, typename Put>
__device__ __forceinline__ void solid_harmonics(T x, T y, T z, Put&& put) {
    const T r2 = (x * x + y * y) + z * z;
    T cm = T(1), sm = T(0);          // (x+iy)^m
    T cm1 = T(0), sm1 = T(0);        // (x+iy)^(m-1)
#pragma unroll
    for (int m = 0; m <= LMAX; ++m) {
        if (m > 0) {
            const T nc = x * cm - y * sm;
            const T ns = x * sm + y * cm;
            cm1 = cm;
            sm1 = sm;
            cm = nc;
            sm = ns;
        }
        T q2 = T(0), q2z = T(0), q2r = T(0);        // Q_{l-2}
        T q1 = T(kDoubleFact[m]), q1z = T(0), q1r = T(0);   // Q_{l-1} (starts at Q_m^m)
#pragma unroll
        for (int l = m; l <= LMAX; ++l) {
            T q, qz, qr;
            if (l == m) {
                q = q1;
                qz = T(0);
                qr = T(0);
            } else if (l == m + 1) {
                q = T(2 * m + 1) * z * q1;
                qz = T(2 * m + 1) * q1;
                qr = T(0);
            } else {
                const T a = T(2 * l - 1), b = T(l + m - 1), inv = T(1.0 / (l - m));
                q = (a * z * q1 - b * r2 * q2) * inv;
                qz = (a * (q1 + z * q1z) - b * r2 * q2z) * inv;
                qr = (a * z * q1r - b * (q2 + r2 * q2r)) * inv;
            }
            if (l > m) {
                q2 = q1; q2z = q1z; q2r = q1r;
                q1 = q;  q1z = qz;  q1r = qr;
            }
            const T K = T(kSHNorm[l][m]);
            const int c0 = l * l + l;
            if (m == 0) {
                T gx = T(0), gy = T(0), gz = T(0);
                if (GRAD) {
                    gx = K * T(2) * x * qr;
                    gy = K * T(2) * y * qr;
                    gz = K * (qz + T(2) * z * qr);
                }
                put(c0, K * q, gx, gy, gz);
            } else {
                T gxp = T(0), gyp = T(0), gzp = T(0), gxm = T(0), gym = T(0), gzm = T(0);
                if (GRAD) {
                    const T dqx = T(2) * x * qr, dqy = T(2) * y * qr, dqz = qz + T(2) * z * qr;
                    const T fm = T(m);
                    gxp = K * (dqx * cm + q * fm * cm1);
                    gyp = K * (dqy * cm - q * fm * sm1);
                    gzp = K * dqz * cm;
                    gxm = K * (dqx * sm + q * fm * sm1);
                    gym = K * (dqy * sm + q * fm * cm1);
                    gzm = K * dqz * sm;
                }
                put(c0 + m, K * q * cm, gxp, gyp, gzp);
                put(c0 - m, K * q * sm, gxm, gym, gzm);
            }
        }
    }
}


// Values only, for the orders m = phase, phase + 4, phase + 8, ... (phase is
// uniform across the wavefront, so four wavefronts split one edge's harmonics
// without divergence).  put(l, m, v) with m signed (-l..l).
template <int LMAX, typename T, typename Put>
__device__ __forceinline__ void solid_harmonics_stripe(T x, T y, T z, int phase, Put&& put) {
    const T r2 = (x * x + y * y) + z * z;
    T cm = T(1), sm = T(0);          // (x+iy)^m
#pragma unroll
    for (int m = 0; m <= LMAX; ++m) {
        if (m > 0) {
            const T nc = x * cm - y * sm;
            const T ns = x * sm + y * cm;
            cm = nc;
            sm = ns;
        }
        if ((m & 3) != phase) continue;
        T q2 = T(0), q1 = T(kDoubleFact[m]);
#pragma unroll
        for (int l = m; l <= LMAX; ++l) {
            T q;
            if (l == m) {
                q = q1;
            } else if (l == m + 1) {
                q = T(2 * m + 1) * z * q1;
            } else {
                q = (T(2 * l - 1) * z * q1 - T(l + m - 1) * r2 * q2) * T(1.0 / (l - m));
            }
            if (l > m) {
                q2 = q1;
                q1 = q;
            }
            const T K = T(kSHNorm[l][m]);
            if (m == 0) {
                put(l, 0, K * q);
            } else {
                put(l, m, K * q * cm);
                put(l, -m, K * q * sm);
            }
        }
    }
}

// l-major variant of the same harmonics: on_l(l) once at the start of degree l,
// then put(l, lm, v, gx, gy, gz) for m = -l..l (lm ascending).  Keeps the degree
// l-1 and l-2 columns of the recurrence for every m in registers.
template <int LMAX, bool GRAD, typename T, typename OnL, typename Put>
__device__ __forceinline__ void solid_harmonics_lmajor(T x, T y, T z, OnL&& on_l, Put&& put) {
    const T r2 = (x * x + y * y) + z * z;
    T cm[LMAX + 1], sm[LMAX + 1];
    cm[0] = T(1);
    sm[0] = T(0);
#pragma unroll
    for (int m = 1; m <= LMAX; ++m) {
        cm[m] = x * cm[m - 1] - y * sm[m - 1];
        sm[m] = x * sm[m - 1] + y * cm[m - 1];
    }
    T q1[LMAX + 1], q1z[LMAX + 1], q1r[LMAX + 1], q2[LMAX + 1], q2z[LMAX + 1], q2r[LMAX + 1];
#pragma unroll
    for (int m = 0; m <= LMAX; ++m) q1[m] = q1z[m] = q1r[m] = q2[m] = q2z[m] = q2r[m] = T(0);
#pragma unroll
    for (int l = 0; l <= LMAX; ++l) {
        T q[LMAX + 1], qz[LMAX + 1], qr[LMAX + 1];
#pragma unroll
        for (int m = 0; m <= l; ++m) {
            if (m == l) {
                q[m] = T(kDoubleFact[m]);
                qz[m] = T(0);
                qr[m] = T(0);
            } else if (m == l - 1) {
                q[m] = T(2 * m + 1) * z * q1[m];
                qz[m] = T(2 * m + 1) * q1[m];
                qr[m] = T(0);
            } else {
                const T a = T(2 * l - 1), bb = T(l + m - 1), inv = T(1.0 / (l - m));
                q[m] = (a * z * q1[m] - bb * r2 * q2[m]) * inv;
                qz[m] = (a * (q1[m] + z * q1z[m]) - bb * r2 * q2z[m]) * inv;
                qr[m] = (a * z * q1r[m] - bb * (q2[m] + r2 * q2r[m])) * inv;
            }
        }
        const int c0 = l * l + l;
        on_l(l);
#pragma unroll
        for (int mm = -l; mm <= l; ++mm) {
            const int am = mm < 0 ? -mm : mm;
            const T K = T(kSHNorm[l][am]);
            T v, gx = T(0), gy = T(0), gz = T(0);
            if (am == 0) {
                v = K * q[0];
                if (GRAD) {
                    gx = K * T(2) * x * qr[0];
                    gy = K * T(2) * y * qr[0];
                    gz = K * (qz[0] + T(2) * z * qr[0]);
                }
            } else {
                const T dqx = T(2) * x * qr[am], dqy = T(2) * y * qr[am], dqz = qz[am] + T(2) * z * qr[am];
                const T fm = T(am);
                if (mm > 0) {
                    v = K * q[am] * cm[am];
                    if (GRAD) {
                        gx = K * (dqx * cm[am] + q[am] * fm * cm[am - 1]);
                        gy = K * (dqy * cm[am] - q[am] * fm * sm[am - 1]);
                        gz = K * dqz * cm[am];
                    }
                } else {
                    v = K * q[am] * sm[am];
                    if (GRAD) {
                        gx = K * (dqx * sm[am] + q[am] * fm * sm[am - 1]);
                        gy = K * (dqy * sm[am] + q[am] * fm * cm[am - 1]);
                        gz = K * dqz * sm[am];
                    }
                }
            }
            put(l, c0 + mm, v, gx, gy, gz);
        }
#pragma unroll
        for (int m = 0; m <= l; ++m) {
            q2[m] = q1[m];
            q2z[m] = q1z[m];
            q2r[m] = q1r[m];
            q1[m] = q[m];
            q1z[m] = qz[m];
            q1r[m] = qr[m];
        }
    }
}

// Adjoint of the solid harmonics S_lm = K_lm Q_lm(z, r^2) Phi_m(x, y), with the
// normalisation K_lm folded into the caller's coefficients.  For each degree l,
// on_l(l) once, then h(lm, h0, h1) returns the caller's two coefficients of lm
// (h0 multiplies grad S, h1 multiplies x S).  Accumulates
//   sum_lm [h1 x S_lm + h0 grad S_lm] / K_lm = x P + (X, Y, Z)
// using grad Q = 2 (x, y, z) dQ/dr^2 + z-hat dQ/dz and
// grad (x+iy)^m = m (x+iy)^(m-1) (1, i): per (l, |m|) pair the two orders share
// Q, dQ/dz, dQ/dr^2, so the gradient costs a few fused products per lm instead
// of forming the three gradient components of every harmonic.
template <int LMAX, typename T, typename OnL, typename H>
__device__ __forceinline__ void solid_harmonics_adjoint(T x, T y, T z, OnL&& on_l, H&& h, T& P, T& X, T& Y, T& Z) {
    const T r2 = (x * x + y * y) + z * z;
    T cm[LMAX + 1], sm[LMAX + 1];
    cm[0] = T(1);
    sm[0] = T(0);
#pragma unroll
    for (int m = 1; m <= LMAX; ++m) {
        cm[m] = x * cm[m - 1] - y * sm[m - 1];
        sm[m] = x * sm[m - 1] + y * cm[m - 1];
    }
    T q1[LMAX + 1], q1z[LMAX + 1], q1r[LMAX + 1], q2[LMAX + 1], q2z[LMAX + 1], q2r[LMAX + 1];
#pragma unroll
    for (int m = 0; m <= LMAX; ++m) q1[m] = q1z[m] = q1r[m] = q2[m] = q2z[m] = q2r[m] = T(0);
#pragma unroll
    for (int l = 0; l <= LMAX; ++l) {
        T q[LMAX + 1], qz[LMAX + 1], qr[LMAX + 1];
#pragma unroll
        for (int m = 0; m <= l; ++m) {
            if (m == l) {
                q[m] = T(kDoubleFact[m]);
                qz[m] = T(0);
                qr[m] = T(0);
            } else if (m == l - 1) {
                q[m] = T(2 * m + 1) * z * q1[m];
                qz[m] = T(2 * m + 1) * q1[m];
                qr[m] = T(0);
            } else {
                const T a = T(2 * l - 1), bb = T(l + m - 1), inv = T(1.0 / (l - m));
                q[m] = (a * z * q1[m] - bb * r2 * q2[m]) * inv;
                qz[m] = (a * (q1[m] + z * q1z[m]) - bb * r2 * q2z[m]) * inv;
                qr[m] = (a * z * q1r[m] - bb * (q2[m] + r2 * q2r[m])) * inv;
            }
        }
        const int c0 = l * l + l;
        on_l(l);
        {
            T h0, h1;
            h(c0, h0, h1);
            P += q[0] * h1 + (T(2) * qr[0]) * h0;
            Z += qz[0] * h0;
        }
#pragma unroll
        for (int m = 1; m <= l; ++m) {
            T h0p, h1p, h0n, h1n;
            h(c0 + m, h0p, h1p);                       // cos-type (x+iy)^m real part
            h(c0 - m, h0n, h1n);                       // sin-type imaginary part
            const T q2r_ = T(2) * qr[m];
            const T tp = q[m] * h1p + q2r_ * h0p, tn = q[m] * h1n + q2r_ * h0n;
            P += cm[m] * tp + sm[m] * tn;
            Z += qz[m] * (cm[m] * h0p + sm[m] * h0n);
            const T u = T(m) * q[m];
            X += u * (cm[m - 1] * h0p + sm[m - 1] * h0n);
            Y += u * (cm[m - 1] * h0n - sm[m - 1] * h0p);
        }
#pragma unroll
        for (int m = 0; m <= l; ++m) {
            q2[m] = q1[m];
            q2z[m] = q1z[m];
            q2r[m] = q1r[m];
            q1[m] = q[m];
            q1z[m] = qz[m];
            q1r[m] = qr[m];
        }
    }
}

// ---------------------------------------------------------------------------
// weighting (w(r) and dw/dr; w0 and zero slope below 1e-12)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void weighting(const cusoap_basis_t& b, double r, double& w, double& dw) {
    if (r < 1e-12) {
        w = b.w0;
        dw = 0.0;
        return;
    }
    switch (b.weighting) {
        case CUSOAP_W_POLY: {
            const double x = r / b.w_r0;
            const double base = 1.0 + 2.0 * (x * x * x) - 3.0 * (x * x);
            if (r <= b.w_r0 && base > 0.0) {
                w = b.w_c * pow(base, b.w_m);
                dw = b.w_c * b.w_m * pow(base, b.w_m - 1.0) * (6.0 * x * x - 6.0 * x) / b.w_r0;
            } else {
                w = 0.0;
                dw = 0.0;
            }
            return;
        }
        case CUSOAP_W_POW: {
            const double x = r / b.w_r0;
            const double xm = pow(x, b.w_m);
            const double den = b.w_d + xm;
            w = b.w_c / den;
            dw = -b.w_c * (b.w_m * pow(x, b.w_m - 1.0) / b.w_r0) / (den * den);
            return;
        }
        case CUSOAP_W_EXP: {
            const double e = exp(-r / b.w_r0);
            const double den = b.w_d + e;
            w = b.w_c / den;
            dw = b.w_c * e / b.w_r0 / (den * den);
            return;
        }
        default:
            w = 1.0;
            dw = 0.0;
            return;
    }
}

// ---------------------------------------------------------------------------
// polynomial basis: I_nl(r), J_nl(r) by the clamped Bessel recurrence on the
// Gauss-Legendre grid (restates _poly_I_J, soap_generator.py:1437-1507)
// ---------------------------------------------------------------------------
// Calls acc(q, l, P) for l = 0..ltop with P = common * ie_l(t_q).
template <typename Acc>
__device__ __forceinline__ void poly_levels(const cusoap_basis_t& b, double r, int ltop, Acc&& acc) {
    const double eta = b.eta;
    for (int q = 0; q < b.quad_n; ++q) {
        const double rx = b.poly_rx[q];
        double tt = (2.0 * eta) * rx * r;
        tt = tt < 1e-12 ? 1e-12 : tt;
        const double e2 = exp(-2.0 * tt);
        const double oOt = 1.0 / tt;
        const double d = rx - r;
        const double common = b.poly_rxw[q] * exp(-eta * (d * d));
        double Ppp = 0.0, Pp = 0.0;
        for (int l = 0; l <= ltop; ++l) {
            double P;
            if (l == 0) {
                P = common * ((1.0 - e2) * (0.5 * oOt));
            } else if (l == 1) {
                double ie1 = (1.0 + e2) * (0.5 * oOt) - (1.0 - e2) * (0.5 * oOt * oOt);
                ie1 = ie1 < 0.0 ? 0.0 : ie1;
                P = common * ie1;
            } else {
                double v = Ppp - (double)(2 * l - 1) * oOt * Pp;
                v = v < 0.0 ? 0.0 : v;
                P = v < Pp ? v : Pp;
            }
            acc(q, l, P);
            Ppp = Pp;
            Pp = P;
        }
    }
}

}  // namespace cusoap
