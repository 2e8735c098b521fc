// hbem_device.hpp -- device-side math shared by the assembly kernels.
//
//  * complex double helpers
//  * J0, Y0, J1, Y1 of real x >= 0 (Hankel H0 = J0 + iY0, H1 = J1 + iY1) with the
//    branch points of the reference's Hankel1Real01::jy_01
//    (cbessel.cpp:1896-1980): x <= 0, x < 1e-8, x < 8, x <= 1e6, x > 1e6
//  * panel evaluation for line / arc / trig-curve panels
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "hankel_tables.inc"

namespace hbem {

struct cdbl {
    double x, y;
};
// single-precision complex: the storage of the mixed-precision LU factors
struct cflt {
    float x, y;
};

__host__ __device__ __forceinline__ cdbl mk(double a, double b) { return cdbl{a, b}; }
__host__ __device__ __forceinline__ cdbl operator+(cdbl a, cdbl b) { return cdbl{a.x + b.x, a.y + b.y}; }
__host__ __device__ __forceinline__ cdbl operator-(cdbl a, cdbl b) { return cdbl{a.x - b.x, a.y - b.y}; }
__host__ __device__ __forceinline__ cdbl operator*(double s, cdbl a) { return cdbl{s * a.x, s * a.y}; }
__host__ __device__ __forceinline__ cdbl operator*(cdbl a, cdbl b) { return cdbl{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ __forceinline__ cdbl operator*(cdbl a, double s) { return cdbl{s * a.x, s * a.y}; }
__host__ __device__ __forceinline__ cdbl cconj(cdbl a) { return cdbl{a.x, -a.y}; }
__host__ __device__ __forceinline__ double cabs2(cdbl a) { return a.x * a.x + a.y * a.y; }
// acc += s * a  (s real)
__device__ __forceinline__ void axpy(cdbl& acc, double s, cdbl a) { acc.x = fma(s, a.x, acc.x); acc.y = fma(s, a.y, acc.y); }

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoOverPi = 0.63661977236758134308;
constexpr double kOneOverPi = 0.31830988618379067154;
constexpr double kEuler = 0.57721566490153286061;
constexpr double kDblEps = 2.220446049250313080847e-16;

// The 128 coefficients of hankel_tables.inc live in constant memory and are
// read (scalar loads, uniform addresses) at the point of use.  As immediates
// they would be loop-invariant, and the compiler keeps loop invariants in
// registers: the coefficients alone would then take every VGPR of the node
// loop.  bessel_table() hides the table address behind an empty asm so the
// loads stay inside the loop.
static __constant__ double kBesselTable[HBEM_TAB_SIZE] = HBEM_BESSEL_TABLE_INIT;

// pointer into the constant address space: uniform reads through it compile
// to scalar loads (s_load), which keep the coefficients in SGPRs
typedef const __attribute__((address_space(4))) double* ctab_ptr;

__device__ __forceinline__ ctab_ptr bessel_table() {
    unsigned long long a = (unsigned long long)(const void*)kBesselTable;
    asm volatile("" : "+s"(a));
    return (ctab_ptr)a;
}

// a * b + c with the wave-uniform c taken straight from its SGPR pair.  Left
// to itself the compiler selects v_fmac_f64, whose addend must sit in the
// destination VGPR, and copies every coefficient from SGPRs to VGPRs first.
__device__ __forceinline__ double fma_su(double a, double b, double c) {
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}

// Four polynomials of degree DEG in one variable, evaluated in lockstep
// (independent Horner chains interleaved); c[4 i + f] is coefficient i of
// polynomial f.  The coefficients of two consecutive steps are read together
// (one 64-byte scalar load, one wait per two steps).
template <int DEG>
__device__ __forceinline__ void horner4(ctab_ptr c, double z, double& v0, double& v1, double& v2, double& v3) {
    double k[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) k[f] = c[4 * DEG - 4 + f];
    v0 = fma_su(z, k[4], k[0]);
    v1 = fma_su(z, k[5], k[1]);
    v2 = fma_su(z, k[6], k[2]);
    v3 = fma_su(z, k[7], k[3]);
    int i = DEG - 2;
#pragma unroll
    for (; i >= 1; i -= 2) {
#pragma unroll
        for (int f = 0; f < 8; ++f) k[f] = c[4 * (i - 1) + f];
        v0 = fma_su(v0, z, k[4]);
        v1 = fma_su(v1, z, k[5]);
        v2 = fma_su(v2, z, k[6]);
        v3 = fma_su(v3, z, k[7]);
        v0 = fma_su(v0, z, k[0]);
        v1 = fma_su(v1, z, k[1]);
        v2 = fma_su(v2, z, k[2]);
        v3 = fma_su(v3, z, k[3]);
    }
    if (i == 0) {
        v0 = fma_su(v0, z, c[0]);
        v1 = fma_su(v1, z, c[1]);
        v2 = fma_su(v2, z, c[2]);
        v3 = fma_su(v3, z, c[3]);
    }
}

// sin x and cos x for 8 <= x <= 1e6: x = n pi/2 + r with n = rint(2x/pi) and r
// from an FMA Cody-Waite reduction against pi/2 split into three doubles;
// sin r = r + r u S(u), cos r = 1 - u/2 + u^2 C(u), u = r^2, with S and C from
// hankel_tables.inc; the quadrant n mod 4 then swaps and negates.
__device__ __forceinline__ void sincos_med(double x, double& s, double& c) {
    const ctab_ptr T = bessel_table() + HBEM_TAB_SC;
    const double n = rint(x * T[0]);
    double r = fma(-n, T[1], x);
    r = fma(-n, T[2], r);
    r = fma(-n, T[3], r);
    const double u = r * r;
    const ctab_ptr k = T + 4;
    double ps = fma_su(u, k[2 * HBEM_SC_DEG], k[2 * HBEM_SC_DEG - 2]);
    double pc = fma_su(u, k[2 * HBEM_SC_DEG + 1], k[2 * HBEM_SC_DEG - 1]);
#pragma unroll
    for (int i = HBEM_SC_DEG - 2; i >= 0; --i) {
        ps = fma_su(ps, u, k[2 * i]);
        pc = fma_su(pc, u, k[2 * i + 1]);
    }
    const double sr = fma(r * u, ps, r);
    const double cr = fma(u * u, pc, fma(-0.5, u, 1.0));
    const int q = (int)n & 3;
    const double s0 = (q & 1) ? cr : sr, c0 = (q & 1) ? sr : cr;
    s = (q & 2) ? -s0 : s0;
    c = ((q + 1) & 2) ? -c0 : c0;
}

// J0, Y0, J1, Y1 of x.  The branch points are the reference's
// (Hankel1Real01::jy_01, cbessel.cpp:1896-1980):
//   x <= 0       J0 = 1, J1 = 0, Y0 = Y1 = 0 (the reference zeroes Y there, cbessel.cpp:1976-1980)
//   x < 8        ascending-series form (A&S 9.1.10-9.1.13) with its four entire
//                parts A, B, C, D from the fits of hankel_tables.inc:
//                J0 = A, J1 = x B, Y0 = (2/pi) ln(x/2) J0 + C,
//                Y1 = (2/pi) ln(x/2) J1 - 2/(pi x) + x D.  Below 1e-8 this
//                agrees with the reference's limiting forms to rounding.
//   x <= 1e6     Hankel asymptotic form with P_nu, x Q_nu from the polynomial
//                fits of hankel_tables.inc (A&S 9.2.5-9.2.6)
//   x > 1e6      leading asymptotic term (cbessel.cpp:1847-1856)
// TINY (touching panel pairs): below x = HBEM_TINY_X the entire parts come
// from the degree-HBEM_TINY_DEG fit of that range.
template <bool TINY = false>
__device__ __forceinline__ void bessel_jy01(double x, double& j0, double& y0, double& j1, double& y1) {
    if (x < 8.0) {
        const double xs = x > 0.0 ? x : 1.0;
        const double h = 0.125 * xs;
        const double y = h * h;
        double A, B, C, D;
        if (TINY && xs < (double)HBEM_TINY_X) horner4<HBEM_TINY_DEG>(bessel_table() + HBEM_TAB_TINY, y, A, B, C, D);
        else horner4<HBEM_SMALL_DEG>(bessel_table() + HBEM_TAB_SMALL, y, A, B, C, D);
        const double L2 = kTwoOverPi * log(0.5 * xs);
        const double xinv = 1.0 / xs;
        const double J1 = xs * B;
        const bool pos = x > 0.0;
        j0 = pos ? A : 1.0;
        j1 = pos ? J1 : 0.0;
        y0 = pos ? fma(L2, A, C) : 0.0;
        y1 = pos ? fma(L2, J1, fma(xs, D, -kTwoOverPi * xinv)) : 0.0;
    } else {
        double s, c;
        if (x <= 1e6) sincos_med(x, s, c);
        else sincos(x, &s, &c);
        const double b = 1.0 / sqrt(kPi * x);      // sqrt(2/(pi x)) / sqrt(2)
        const double cs = c + s, sc = s - c;       // sqrt2 cos(x - pi/4), sqrt2 sin(x - pi/4)
        if (x <= 1e6) {
            const double xinv = 1.0 / x;
            const double z = 64.0 * xinv * xinv;
            double P0, Q0, P1, Q1;
            horner4<HBEM_PQ_DEG>(bessel_table() + HBEM_TAB_PQ, z, P0, Q0, P1, Q1);
            Q0 *= xinv;
            Q1 *= xinv;
            j0 = b * fma(P0, cs, -Q0 * sc);
            y0 = b * fma(P0, sc, Q0 * cs);
            j1 = b * fma(P1, sc, Q1 * cs);
            y1 = b * fma(Q1, sc, -P1 * cs);
        } else {
            j0 = b * cs; y0 = b * sc; j1 = b * sc; y1 = -b * cs;
        }
    }
}

// ---------------------------------------------------------------------------
// Complex helpers and H0, H1 of complex argument (wavenumbers that are not
// real positive: GalerkinBuilder::H1_01_cplx, galerkin_builder.cpp:224-235)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double recip(double a) { return 1.0 / a; }
__device__ __forceinline__ cdbl recip(cdbl a) {
    const double s = 1.0 / (a.x * a.x + a.y * a.y);
    return mk(a.x * s, -a.y * s);
}
__device__ __forceinline__ cdbl cdiv(cdbl a, cdbl b) { return a * recip(b); }
__device__ __forceinline__ double cabsd(cdbl a) { return hypot(a.x, a.y); }
__device__ __forceinline__ cdbl clogd(cdbl a) { return mk(log(cabsd(a)), atan2(a.y, a.x)); }
__device__ __forceinline__ cdbl cexpd(cdbl a) {
    double s, c;
    sincos(a.y, &s, &c);
    const double e = exp(a.x);
    return mk(e * c, e * s);
}
__device__ __forceinline__ cdbl csqrtd(cdbl a) {   // principal branch
    const double r = cabsd(a);
    if (r == 0.0) return mk(0.0, 0.0);
    const double t = sqrt(0.5 * (r + fabs(a.x)));
    if (a.x >= 0.0) return mk(t, 0.5 * a.y / t);
    return mk(0.5 * fabs(a.y) / t, a.y < 0.0 ? -t : t);
}

// K0(w), K1(w) for Re w >= 0, w != 0: the ascending series (A&S 9.6.10-9.6.13)
// for |w| <= 2, Temme's continued fraction evaluated by Steed's algorithm
// (Thompson & Barnett 1986) above; the loops end on convergence and are
// bounded (a non-finite input runs to the bound and yields non-finite values).
__device__ inline void bessel_k01_cplx(cdbl w, cdbl& k0, cdbl& k1) {
    if (cabsd(w) <= 2.0) {
        const cdbl t = 0.25 * (w * w), lw = clogd(0.5 * w);
        cdbl term = mk(1.0, 0.0), u = mk(1.0, 0.0), i0 = mk(1.0, 0.0), s0 = mk(0.0, 0.0), i1s = mk(1.0, 0.0);
        double H = 0.0;
        cdbl s1 = mk(1.0 - 2.0 * kEuler, 0.0);
        for (int k = 1; k < 200; ++k) {
            term = term * t * (1.0 / double(k * k));
            u = u * t * (1.0 / double(k * (k + 1)));
            const double Hk = H + 1.0 / k, Hk1 = Hk + 1.0 / (k + 1);
            i0 = i0 + term;
            s0 = s0 + Hk * term;
            i1s = i1s + u;
            s1 = s1 + (Hk + Hk1 - 2.0 * kEuler) * u;
            H = Hk;
            if (cabsd(term) < 1e-17 * cabsd(i0) && cabsd(u) < 1e-17 * cabsd(i1s)) break;
        }
        const cdbl l = lw + mk(kEuler, 0.0);
        k0 = s0 - l * i0;
        const cdbl i1 = (0.5 * w) * i1s;
        k1 = recip(w) + lw * i1 - (0.25 * w) * s1;
        return;
    }
    cdbl b = 2.0 * (mk(1.0, 0.0) + w), d = recip(b), h = d, dh = d;
    cdbl q0 = mk(0.0, 0.0), q1 = mk(1.0, 0.0);
    cdbl Q = mk(0.25, 0.0), C = mk(0.25, 0.0), S = mk(1.0, 0.0) + Q * dh;
    double A = -0.25;
    for (int i = 1; i < 1000; ++i) {
        A -= 2.0 * i;
        C = (-A / (i + 1.0)) * C;
        const cdbl qn = (q0 - b * q1) * (1.0 / A);
        q0 = q1;
        q1 = qn;
        Q = Q + C * qn;
        b = b + mk(2.0, 0.0);
        d = recip(b + A * d);
        dh = (b * d - mk(1.0, 0.0)) * dh;
        h = h + dh;
        const cdbl dS = Q * dh;
        S = S + dS;
        if (cabsd(dS) < 1e-17 * cabsd(S)) break;
    }
    h = 0.25 * h;
    k0 = csqrtd(cdiv(mk(kPi, 0.0), 2.0 * w)) * cdiv(cexpd(mk(-w.x, -w.y)), S);
    k1 = cdiv(k0 * (w + mk(0.5, 0.0) - h), w);
}

// H0(z), H1(z) for Im z >= 0: -(2i/pi) K0(-iz), -(2/pi) K1(-iz) (A&S 9.6.4);
// z = 0 and other non-finite results give 0 (galerkin_builder.cpp:216-222)
__device__ inline void hankel01_cplx(cdbl z, cdbl& h0, cdbl& h1) {
    if (z.x == 0.0 && z.y == 0.0) { h0 = mk(0.0, 0.0); h1 = mk(0.0, 0.0); return; }
    cdbl k0, k1;
    bessel_k01_cplx(mk(z.y, -z.x), k0, k1);
    h0 = mk(kTwoOverPi * k0.y, -kTwoOverPi * k0.x);
    h1 = (-kTwoOverPi) * k1;
    if (!(isfinite(h0.x) && isfinite(h0.y))) h0 = mk(0.0, 0.0);
    if (!(isfinite(h1.x) && isfinite(h1.y))) h1 = mk(0.0, 0.0);
}

// ---------------------------------------------------------------------------
// Panels (8 doubles {type, a0..a6}).  Returns the point and the forward
// tangent gamma'(param) at param = s (or at 1 - s when `swapped`), i.e.
// operator[] / Derivative_01, resp. swapped_op / Derivative_01_swapped(neg=true)
// (parametrized_line.cpp:43-107, parametrized_circular_arc.cpp:54-121).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void panel_eval(const double* __restrict__ p, const double* __restrict__ coefs, double s,
                                           bool swapped, cdbl& z, cdbl& tau) {
    const int type = (int)p[0];
    if (type == 0) {
        const double dx = p[3] - p[1], dy = p[4] - p[2];
        if (!swapped) { z = mk(fma(dx, s, p[1]), fma(dy, s, p[2])); }
        else { z = mk(fma(-dx, s, p[3]), fma(-dy, s, p[4])); }
        tau = mk(dx, dy);
    } else if (type == 1) {
        const double d = p[5] - p[4];
        const double th = swapped ? fma(s, -d, p[5]) : fma(s, d, p[4]);
        double sn, cs;
        sincos(th, &sn, &cs);
        z = mk(fma(p[3], cs, p[1]), fma(p[3], sn, p[2]));
        const double a = p[3] * d;
        tau = mk(-a * sn, a * cs);
    } else if (type == 3) {
        // spline piece: rows of 9 doubles per spline piece j from row p[3]:
        // {t_j, x_j, y_j, bx, cx, dx, by, cy, dy}; the piece holding t by bisection
        const double dt = p[2] - p[1];
        const double t = swapped ? fma(s, -dt, p[2]) : fma(s, dt, p[1]);
        const double* c = coefs + 3 * (int)p[3];
        int lo = 0, hi = (int)p[4];
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (c[9 * mid] <= t) lo = mid;
            else hi = mid;
        }
        const double* r = c + 9 * lo;
        const double d = t - r[0];
        z = mk(r[1] + d * (r[3] + d * (r[4] + d * r[5])), r[2] + d * (r[6] + d * (r[7] + d * r[8])));
        tau = mk(dt * (r[3] + d * (2.0 * r[4] + 3.0 * d * r[5])), dt * (r[6] + d * (2.0 * r[7] + 3.0 * d * r[8])));
    } else {
        const double dt = p[2] - p[1];
        const double t = swapped ? fma(s, -dt, p[2]) : fma(s, dt, p[1]);
        const int c0 = (int)p[3], nc = (int)p[4];
        cdbl zz = mk(0.0, 0.0), dz = mk(0.0, 0.0);
        for (int k = 0; k < nc; ++k) {
            const double m = coefs[3 * (c0 + k)];
            const cdbl cm = mk(coefs[3 * (c0 + k) + 1], coefs[3 * (c0 + k) + 2]);
            double sn, cs;
            sincos(m * t, &sn, &cs);
            const cdbl e = cm * mk(cs, sn);
            zz = zz + e;
            dz = dz + mk(-m * e.y, m * e.x);
        }
        z = zz;
        tau = dt * dz;
    }
}

}  // namespace hbem
