// Point-wise numerics of the 2D Euler finite-volume step, shared by the HIP kernels and host code.
//
// Every expression keeps the reference's operation order (binary64, no contraction: the library
// is compiled with -ffp-contract=off) so that results are bit-identical to the reference solver:
//   state2d.hpp:16-36    to_conserved, to_primitive, sound_speed, flux_x
//   limiters.hpp:7-18    none, unlimited, minmod, van Leer
//   solver2d.hpp:17-27   limited slope, face-side value W +- dW/2
//   riemann2d.hpp:11-44  Rusanov; HLLC with Davis speeds; y faces by swapping the momenta
//
// Design: a face flux needs, for each side, the face value W +- dW/2 turned into conserved variables
// and back into primitives plus a sound speed.  That chain depends on one cell only, so it is done
// once per cell edge (`Side`), in the face-normal frame, and the Riemann solver then only combines
// two precomputed sides.  For a y face the normal frame is the x frame with the momenta and the
// velocities relabelled; the relabelling is exact because (u*u + v*v) == (v*v + u*u) in IEEE
// arithmetic, which is all the reference's swap_xy changes.
//
// Arithmetic policy: by default every operation is the IEEE one the reference performs (the `exact`
// namespace).  A translation unit that defines NPDE_FAST_ARITH before including this header gets the
// `fast` namespace instead: divisions by a reciprocal refined by two Newton steps without the final
// rounding correction, a square root without its last correction step, and (with -ffp-contract=fast
// on that unit) fused multiply-adds.  Each such operation is within ~2 ulp of the IEEE result, so a
// step differs from the exact one by rounding-level amounts; it is a separate, opt-in variant.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NPDE_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define NPDE_HD inline
#endif

#if defined(NPDE_FAST_ARITH)
#define NPDE_ARITH_NS fast
#else
#define NPDE_ARITH_NS exact
#endif

namespace npde {
inline namespace NPDE_ARITH_NS {

constexpr double kGamma = 1.4;
constexpr double kGm1 = kGamma - 1.0;  // 0.3999999999999999 (1.4 is not exact in binary64)

enum Scheme : int { kRusanov = 0, kHllc = 1 };
enum LimiterKind : int { kLimNone = 0, kLimUnlimited = 1, kLimMinmod = 2, kLimVanLeer = 3 };
enum Bc : int { kTransmissive = 0, kPeriodic = 1 };

struct Cons { double d, m, n, e; };  // rho, rho*u, rho*v, E
struct Prim { double d, u, v, p; };  // rho, u, v, p

// One side of a face in the face-normal frame: conserved (d, mn, mt, e), the primitives recovered
// from them (un, ut, p), the sound speed a, and the reciprocal rd of d used by its divisions.
struct Side { double d, mn, mt, e, un, ut, p, a, rd; };
struct Flux { double d, mn, mt, e; };

NPDE_HD double dsqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_sqrt(x);
#else
    return std::sqrt(x);
#endif
}
NPDE_HD double dabs(double x) { return __builtin_fabs(x); }

// std::min / std::max semantics: the first argument wins ties and unordered comparisons
NPDE_HD double tmin(double a, double b) { return (b < a) ? b : a; }
NPDE_HD double tmax(double a, double b) { return (a < b) ? b : a; }

// ---- correctly rounded division with a shared reciprocal -----------------------------------
// gfx950's binary64 division x / y is: r = rcp(y) refined by two Newton steps (4 FMAs), q0 = x*r,
// one FMA remainder and one FMA correction, wrapped in div_scale / div_fmas / div_fixup, which only
// rescale operands whose exponents are near the ends of the range and patch zeros, infinities and
// NaNs.  Several of the scheme's divisions share a denominator (rho in to_primitive and in the sound
// speed; the constant gamma - 1 in to_conserved), so the reciprocal part is run once per
// denominator (`recip`) and each quotient costs three instructions (`div_r`).  For operands where
// the rescaling does not trigger -- |x| and |y| in [2^-900, 2^700] or x = 0, i.e. every value a
// physical state produces -- this is the compiler's own instruction sequence, hence the same IEEE
// quotient; the remainder is taken as y*q0 - x so that x = -0 still gives -0.
// Define NPDE_IEEE_DIV to use the plain operator everywhere.
#if defined(__HIP_DEVICE_COMPILE__) && defined(NPDE_FAST_ARITH)
NPDE_HD double recip(double y) {
    double r = __builtin_amdgcn_rcp(y);
    double e = __builtin_fma(-y, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-y, r, 1.0);
    return __builtin_fma(r, e, r);
}
NPDE_HD double div_r(double x, double, double r) { return x * r; }
NPDE_HD double ddiv(double x, double y) { return x * recip(y); }
#elif defined(__HIP_DEVICE_COMPILE__) && !defined(NPDE_IEEE_DIV)
NPDE_HD double recip(double y) {
    double r = __builtin_amdgcn_rcp(y);
    double e = __builtin_fma(-y, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-y, r, 1.0);
    return __builtin_fma(r, e, r);
}
NPDE_HD double div_r(double x, double y, double r) {
    const double q0 = x * r;
    const double rem = __builtin_fma(y, q0, -x);
    return __builtin_fma(-rem, r, q0);
}
NPDE_HD double ddiv(double x, double y) { return x / y; }
#else
NPDE_HD double recip(double y) { return y; }  // the plain path divides by y itself
NPDE_HD double div_r(double x, double y, double) { return x / y; }
NPDE_HD double ddiv(double x, double y) { return x / y; }
#endif

// reciprocal of gamma - 1 for div_r(x, kGm1, rgm1()); a pure function, hoisted by the compiler
NPDE_HD double rgm1() { return recip(kGm1); }

// Square root: gfx950's correctly rounded binary64 sqrt is rsq(x) refined by a Goldschmidt step and two
// FMA corrections, inside a pre/post scaling by 2^256/2^-128 for x < 2^-767 and a fix-up that returns
// x itself for +-0 and +inf.  The sound speed's argument gamma*p/rho is never that small for a state
// with p > 0, so the core runs without the scaling (the fix-up is kept); for x >= 2^-767, zeros,
// infinities, negatives and NaNs this is the compiler's sequence, hence the IEEE result.
// The fast variant stops before the last correction step.
#if defined(__HIP_DEVICE_COMPILE__) && defined(NPDE_FAST_ARITH)
NPDE_HD double sqrt_core(double x) {
    const double g = __builtin_amdgcn_rsq(x);
    double s = x * g;
    double h = g * 0.5;
    const double r = __builtin_fma(-h, s, 0.5);
    s = __builtin_fma(s, r, s);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-s, s, x);
    s = __builtin_fma(d, h, s);
    return __builtin_amdgcn_class(x, 0x260 /* -0, +0, +inf */) ? x : s;
}
#elif defined(__HIP_DEVICE_COMPILE__) && !defined(NPDE_IEEE_DIV)
NPDE_HD double sqrt_core(double x) {
    const double g = __builtin_amdgcn_rsq(x);
    double s = x * g;
    double h = g * 0.5;
    const double r = __builtin_fma(-h, s, 0.5);
    s = __builtin_fma(s, r, s);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-s, s, x);
    s = __builtin_fma(d, h, s);
    d = __builtin_fma(-s, s, x);
    s = __builtin_fma(d, h, s);
    return __builtin_amdgcn_class(x, 0x260 /* -0, +0, +inf */) ? x : s;
}
#else
NPDE_HD double sqrt_core(double x) { return dsqrt(x); }
#endif

NPDE_HD Prim prim_of(double d, double m, double n, double e, double rd) {
    const double u = div_r(m, d, rd);
    const double v = div_r(n, d, rd);
    const double kin = (0.5 * d) * (u * u + v * v);
    return {d, u, v, kGm1 * (e - kin)};
}

NPDE_HD Prim prim_of(double d, double m, double n, double e) { return prim_of(d, m, n, e, recip(d)); }

NPDE_HD Cons cons_of(const Prim& w) {
    const double kin = (0.5 * w.d) * (w.u * w.u + w.v * w.v);
    return {w.d, w.d * w.u, w.d * w.v, div_r(w.p, kGm1, rgm1()) + kin};
}

NPDE_HD double sound(double d, double p, double rd) { return sqrt_core(div_r(kGamma * p, d, rd)); }
NPDE_HD double sound(double d, double p) { return sound(d, p, recip(d)); }

template <int LIM>
NPDE_HD double limit(double a, double b) {
    if constexpr (LIM == kLimNone) {
        return 0.0;
    } else if constexpr (LIM == kLimUnlimited) {
        return 0.5 * (a + b);
    } else if constexpr (LIM == kLimMinmod) {
        const double s = (dabs(a) < dabs(b)) ? a : b;
        return (a * b <= 0.0) ? 0.0 : s;
    } else {
        const double s = ddiv(2.0 * a * b, a + b);  // harmonic mean; unused where a*b <= 0
        return (a * b <= 0.0) ? 0.0 : s;
    }
}

template <int LIM>
NPDE_HD Prim slope(const Prim& m, const Prim& c, const Prim& p) {
    return {limit<LIM>(c.d - m.d, p.d - c.d), limit<LIM>(c.u - m.u, p.u - c.u),
            limit<LIM>(c.v - m.v, p.v - c.v), limit<LIM>(c.p - m.p, p.p - c.p)};
}

// Side state of the cell edge at W + h*dW (h = +0.5 right/top edge, -0.5 left/bottom edge),
// returned in the x frame (mn = rho*u, un = u).  to_conserved then to_primitive, as the reference's
// flux call does on edge values.
NPDE_HD Side edge_side(const Prim& w, const Prim& dw, double h) {
    const Prim we{w.d + h * dw.d, w.u + h * dw.u, w.v + h * dw.v, w.p + h * dw.p};
    const Cons c = cons_of(we);
    const double rd = recip(c.d);
    const Prim r = prim_of(c.d, c.m, c.n, c.e, rd);
    return {c.d, c.m, c.n, c.e, r.u, r.v, r.p, sound(r.d, r.p, rd), rd};
}

// x frame -> y frame (normal = y): relabel momenta and velocities
NPDE_HD Side to_y_frame(const Side& s) { return {s.d, s.mt, s.mn, s.e, s.ut, s.un, s.p, s.a, s.rd}; }

NPDE_HD Flux phys_flux(const Side& s) {
    return {s.mn, s.mn * s.un + s.p, s.mt * s.un, s.un * (s.e + s.p)};
}

NPDE_HD Side pick(bool first, const Side& a, const Side& b) {
    return {first ? a.d : b.d,   first ? a.mn : b.mn, first ? a.mt : b.mt, first ? a.e : b.e,
            first ? a.un : b.un, first ? a.ut : b.ut, first ? a.p : b.p,   first ? a.a : b.a,
            first ? a.rd : b.rd};
}

// HLLC with Davis wave speeds (riemann2d.hpp:18-39).  The reference's four-way branch
//   0<=SL: F(UL) | 0<=S*: F(UL)+SL(UL*-UL) | 0<=SR: F(UR)+SR(UR*-UR) | else F(UR)
// is evaluated branch-free: one side is selected, its star state computed once, and the star
// correction kept or dropped.  Each kept value is produced by the same operations as the reference's.
NPDE_HD Flux hllc(const Side& A, const Side& B) {
    const double sl = tmin(A.un - A.a, B.un - B.a);
    const double sr = tmax(A.un + A.a, B.un + B.a);
    const double num = ((B.p - A.p) + (A.d * A.un) * (sl - A.un)) - (B.d * B.un) * (sr - B.un);
    const double den = A.d * (sl - A.un) - B.d * (sr - B.un);
    const double ss = ddiv(num, den);
    const bool left = (0.0 <= sl) || (0.0 <= ss);
    const bool star = left ? !(0.0 <= sl) : (0.0 <= sr);
    const Side X = pick(left, A, B);
    const double S = left ? sl : sr;
    const Flux F = phys_flux(X);
    const double q = X.d * (S - X.un);
    const double f = ddiv(q, S - ss);
    const double sd = f, smn = f * ss, smt = f * X.ut;
    const double se = f * (div_r(X.e, X.d, X.rd) + (ss - X.un) * (ss + ddiv(X.p, q)));
    const Flux Fs{F.d + S * (sd - X.d), F.mn + S * (smn - X.mn), F.mt + S * (smt - X.mt),
                  F.e + S * (se - X.e)};
    return {star ? Fs.d : F.d, star ? Fs.mn : F.mn, star ? Fs.mt : F.mt, star ? Fs.e : F.e};
}

// Rusanov (riemann2d.hpp:11-15)
NPDE_HD Flux rusanov(const Side& A, const Side& B) {
    const double s = tmax(dabs(A.un) + A.a, dabs(B.un) + B.a);
    const Flux FA = phys_flux(A), FB = phys_flux(B);
    const double h = 0.5 * s;
    return {0.5 * (FA.d + FB.d) - h * (B.d - A.d), 0.5 * (FA.mn + FB.mn) - h * (B.mn - A.mn),
            0.5 * (FA.mt + FB.mt) - h * (B.mt - A.mt), 0.5 * (FA.e + FB.e) - h * (B.e - A.e)};
}

template <int SCHEME>
NPDE_HD Flux riemann(const Side& A, const Side& B) {
    if constexpr (SCHEME == kRusanov) return rusanov(A, B);
    else return hllc(A, B);
}

// y-face flux from two x-frame sides (bottom A, top B); result back in (rho, mx, my, E) order
template <int SCHEME>
NPDE_HD Cons flux_y(const Side& A, const Side& B) {
    const Flux f = riemann<SCHEME>(to_y_frame(A), to_y_frame(B));
    return {f.d, f.mt, f.mn, f.e};
}

template <int SCHEME>
NPDE_HD Cons flux_x(const Side& A, const Side& B) {
    const Flux f = riemann<SCHEME>(A, B);
    return {f.d, f.mn, f.mt, f.e};
}

// Flux scale of a face for the learned increments (no counterpart in the reference): r a, r a^2, r a^2,
// r a^3 with a the larger sound speed of the two cells sharing the face and r = d^2 / (d + kCorrRho) for
// d their smaller density (arguments: the cell left of / below the face, then the one right of / above
// it).  The increment added to the face flux is the network's dimensionless factor times this scale,
// evaluated on each stage's input state, so a correction stays a bounded fraction of the local flux as
// the flow evolves.  r ~ d for d >> kCorrRho; below it the corrections fade out (r ~ d^2 / kCorrRho):
// toward vacuum at finite pressure the sound speed, and with it the scale, grows without bound, and a
// correction that removed a small fraction of a cell's mass per step could drain it step by step.
constexpr double kCorrRho = 0.05;
NPDE_HD Cons corr_scale(double dl, double al, double dr, double ar) {
    const double d = tmin(dl, dr), a = tmax(al, ar);
    const double r = ddiv(d * d, d + kCorrRho);
    const double s1 = r * a, s2 = s1 * a;
    return {s1, s2, s2, s2 * a};
}

// F + k * s, component-wise (a face flux plus its scaled increment)
NPDE_HD Cons add_scaled(const Cons& F, const Cons& k, const Cons& s) {
    return {F.d + k.d * s.d, F.m + k.m * s.m, F.n + k.n * s.n, F.e + k.e * s.e};
}

// Ghost index -> interior index it is copied from (grid2d.hpp:43-66); corners are never needed.
NPDE_HD int remap(int i, int n, int bc) {
    if (i >= 0 && i < n) return i;
    if (bc == kPeriodic) {
        int r = i % n;
        return r < 0 ? r + n : r;
    }
    return i < 0 ? 0 : n - 1;
}

}  // namespace NPDE_ARITH_NS
}  // namespace npde
