// Shared device/host definitions for the gfx950 DBS kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace pngf {

typedef double2 zd;

__host__ __device__ inline zd zmk(double x, double y) { zd r; r.x = x; r.y = y; return r; }
__host__ __device__ inline zd zadd(zd a, zd b) { return zmk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline zd zsub(zd a, zd b) { return zmk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline zd zmul(zd a, zd b) { return zmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline zd zscale(zd a, double s) { return zmk(a.x * s, a.y * s); }
// acc + a*b
__host__ __device__ inline zd zfma(zd a, zd b, zd acc) {
    return zmk(acc.x + a.x * b.x - a.y * b.y, acc.y + a.x * b.y + a.y * b.x);
}
// acc - a*b
__host__ __device__ inline zd zfms(zd a, zd b, zd acc) {
    return zmk(acc.x - a.x * b.x + a.y * b.y, acc.y - a.x * b.y - a.y * b.x);
}
__host__ __device__ inline zd zdiv(zd a, zd b) {
    double d = b.x * b.x + b.y * b.y;
    return zmk((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__host__ __device__ inline zd zrecip(zd b) {
    double d = b.x * b.x + b.y * b.y;
    return zmk(b.x / d, -b.y / d);
}
__host__ __device__ inline double zabs1(zd a) { return fabs(a.x) + fabs(a.y); }
__host__ __device__ inline double zabs2(zd a) { return a.x * a.x + a.y * a.y; }

// complex64: the state and arithmetic of the single-precision mode (the reference's
// USE_SINGLE_PRECISION build, pngf-opt/main.cpp:58-72), same operations as the zd ones
typedef float2 zf;
__host__ __device__ inline zf zmkf(float x, float y) { zf r; r.x = x; r.y = y; return r; }
__host__ __device__ inline zf zadd(zf a, zf b) { return zmkf(a.x + b.x, a.y + b.y); }
__host__ __device__ inline zf zsub(zf a, zf b) { return zmkf(a.x - b.x, a.y - b.y); }
__host__ __device__ inline zf zmul(zf a, zf b) { return zmkf(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline zf zscale(zf a, float s) { return zmkf(a.x * s, a.y * s); }
__host__ __device__ inline zf zfma(zf a, zf b, zf acc) {
    return zmkf(acc.x + a.x * b.x - a.y * b.y, acc.y + a.x * b.y + a.y * b.x);
}
__host__ __device__ inline zf zfms(zf a, zf b, zf acc) {
    return zmkf(acc.x - a.x * b.x + a.y * b.y, acc.y - a.x * b.y - a.y * b.x);
}
__host__ __device__ inline zf zdiv(zf a, zf b) {
    float d = b.x * b.x + b.y * b.y;
    return zmkf((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__host__ __device__ inline zf zrecip(zf b) {
    float d = b.x * b.x + b.y * b.y;
    return zmkf(b.x / d, -b.y / d);
}
__host__ __device__ inline float zabs1(zf a) { return fabsf(a.x) + fabsf(a.y); }
__host__ __device__ inline float zabs2(zf a) { return a.x * a.x + a.y * a.y; }

// The real type of a complex type, and a constructor usable in code generic over both.
template <class Z> struct zreal;
template <> struct zreal<zd> { typedef double type; };
template <> struct zreal<zf> { typedef float type; };
template <class Z>
__host__ __device__ inline Z zc(typename zreal<Z>::type x, typename zreal<Z>::type y) {
    Z r;
    r.x = x;
    r.y = y;
    return r;
}
template <class Z> __host__ __device__ inline Z zneg(Z a) { return zc<Z>(-a.x, -a.y); }
__host__ __device__ inline zd zwide(zd a) { return a; }
__host__ __device__ inline zd zwide(zf a) { return zmk(a.x, a.y); }

// Per-frequency constants read by the objective.
struct FreqConst {
    zd vinc, ln1, ln2;
    double k0;      // 2 pi f / c0
};

// Objective parameters (main.cpp:157-164, 277-289).
struct ObjParams {
    int mode;       // 0 REF, 1 S11
    int nf, ne;
    double zfac;    // dx * 0.5 / (2 dy dz)
    double pfac;    // dx * 0.5 * (2 dy dz) * 0.5
    double z0;
};

// Flip-group tables on the device.
struct GroupTab {
    const int32_t* tiles;   // [n][4]
    const int32_t* ntiles;  // [n]
    const int32_t* comps;   // [n][gmax]
    const int32_t* ncomp;   // [n]
    int gmax;
};

constexpr double kC0 = 2.997924579998211E8;
constexpr double kEta0 = 3.767303134118051e+02;
constexpr double kPi = 3.14159265358979323846;

// S11 at frequency f from that frequency's newEobj E (main.cpp:402-403) and, for mode 0, the
// directivity D (main.cpp:404-406; 0 in mode 1).
__host__ __device__ inline zd s11_of(const ObjParams& op, const FreqConst* fc, const zd* E, int f, double* D) {
    zd u = zadd(E[0], fc[f].vinc);
    zd zin = zmk(-u.x * op.zfac, -u.y * op.zfac);
    zd s11 = zdiv(zmk(zin.x - op.z0, zin.y), zmk(zin.x + op.z0, zin.y));
    *D = 0.0;
    if (op.mode == 0) {
        double k0 = fc[f].k0;
        zd a = zadd(E[1], fc[f].ln1), b2 = zadd(E[2], fc[f].ln2);
        double dd = (k0 * k0 / (32 * kPi * kPi * kEta0)) * (zabs2(a) + zabs2(b2));
        double pin = -u.x * op.pfac;
        *D = 4.0 * kPi * dd / pin;
    }
    return s11;
}

// Objective of one candidate from its newEobj at every frequency, in frequency
// order (main.cpp:401-423): Zin, S11 and, for mode 0, the directivity D.
__host__ __device__ inline double objective_of(const ObjParams& op, const FreqConst* fc, const zd* e) {
    double acc = 0.0;
    for (int f = 0; f < op.nf; f++) {
        double D;
        const zd s11 = s11_of(op, fc, e + f * op.ne, f, &D);
        double m2 = zabs2(s11);
        if (op.mode == 0) {
            double m = sqrt(m2);
            acc += 200.0 * m * m + (15.0 - D) * (15.0 - D);
        } else {
            acc += m2;
        }
    }
    return op.mode == 0 ? acc : acc / op.nf;
}


// Synthetic G entry: dyadic Green's function of a homogeneous medium, softened
// (R^2 -> R^2 + a^2), times -i:  g(R) [ t1 (u_m.u_n) + t2 (u_m.R)(u_n.R) / R^2 ],
// g = exp(-i k R) / (4 pi R), t1 = 1 - i/(kR) - 1/(kR)^2, t2 = -1 + 3i/(kR) + 3/(kR)^2.
// p = (x, y, orientation) in metres, orientation 0 = x, 1 = y.
__host__ __device__ inline zd synth_entry(const double* pm, const double* pn, double kr, double soft2) {
    const double rx = pm[0] - pn[0], ry = pm[1] - pn[1];
    const double R2 = rx * rx + ry * ry + soft2;
    const double R = sqrt(R2);
    const double kR = kr * R;
    const double inv = 1.0 / kR;
    const double c = cos(kR), s = sin(kR);
    const double gm = 1.0 / (4.0 * kPi * R);
    // g = gm (c - i s)
    const zd t1 = zmk(1.0 - inv * inv, -inv);
    const zd t2 = zmk(-1.0 + 3.0 * inv * inv, 3.0 * inv);
    const double um = pm[2] == 0.0 ? 1.0 : 0.0, vm = 1.0 - um;
    const double un = pn[2] == 0.0 ? 1.0 : 0.0, vn = 1.0 - un;
    const double uu = um * un + vm * vn;
    const double ur = (um * rx + vm * ry) * (un * rx + vn * ry) / R2;
    const zd br = zmk(t1.x * uu + t2.x * ur, t1.y * uu + t2.y * ur);
    const zd g = zmk(gm * c, -gm * s);
    const zd v = zmul(g, br);
    return zmk(v.y, -v.x);  // times -i
}

}  // namespace pngf
