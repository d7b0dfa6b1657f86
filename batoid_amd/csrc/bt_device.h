// bt_device.h -- per-ray FP64 device functions of the trace hot path.
//
// Surface sag / normal / time-to-intersect, media n(lambda) and obscuration
// programs, evaluated on descriptors that live in device memory.  All
// threads of a launch read the same descriptor, so its fields are
// wave-uniform (scalar loads); only the per-ray arithmetic is vector work.
//
// Operation order follows the reference expression by expression and the
// library is compiled with -ffp-contract=off, so with IEEE-correct FP64
// division and square root the results are bit-identical to the reference's
// CPU build (which has no FMA on baseline x86-64).  Citations are to
// jmeyers314/batoid.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "../../include/batoid_amd.h"
#include "bt_mlp.h"

#define BT_DEV __device__ __forceinline__

namespace bt {

// ---------------------------------------------------------------------------
// sqrt(x) and 1/sqrt(x), correctly rounded, without range handling where it
// cannot change a bit.
//
// The compiler's IEEE sqrt scales x < 2^-767 by 2^256 before the v_rsq_f64
// estimate, refines it (g = x r, h = r/2; one coupled Goldschmidt step; two
// Newton corrections from the residual d = x - g*g), unscales, and selects x
// itself for +-0 and +inf.  Its IEEE division rescales operands of extreme
// exponent (v_div_scale), refines v_rcp_f64 by two Newton steps, applies one
// Markstein correction q + (n - d q) r, and patches special operands
// (v_div_fixup).  For x in [2^-767, DBL_MAX] all of that range handling is
// the identity: the scale factors are 2^0 and nothing is special, and the
// root s lies in [2^-384, 2^512], where 1/s is not rescaled or patched
// either (the quotient 1*r of the first step is r itself).  There the same
// refinement steps run here without it; the range handling is 8 of the 18
// vector instructions of a sqrt and 12 of the 29 of a 1/sqrt, the range
// test that replaces it two compares.  Every other x (zero, tiny, negative,
// infinite, NaN) takes the compiler's full sequence, so cr_sqrt(x) and
// cr_rsqrt(x) are bit-identical to sqrt(x) and 1/sqrt(x) for every input
// (bt_check_sqrt exposes both for testing).
// ---------------------------------------------------------------------------
BT_DEV bool sqrt_in_plain_range(double x)
{
    return x >= 0x1p-767 && x <= 0x1.fffffffffffffp+1023;
}

BT_DEV double sqrt_plain_range(double x)
{
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r;
    double h = r * 0.5;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

BT_DEV double recip_plain_range(double s)
{
    double r = __builtin_amdgcn_rcp(s);
    double e = __builtin_fma(-s, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-s, r, 1.0);
    r = __builtin_fma(r, e, r);
    const double rem = __builtin_fma(-s, r, 1.0);
    return __builtin_fma(rem, r, r);
}

BT_DEV double cr_sqrt(double x)
{
    if (sqrt_in_plain_range(x)) return sqrt_plain_range(x);
    return __builtin_sqrt(x);
}

BT_DEV double cr_rsqrt(double x)                    // 1/sqrt(x): two roundings
{
    if (sqrt_in_plain_range(x)) return recip_plain_range(sqrt_plain_range(x));
    return 1.0 / __builtin_sqrt(x);
}

// ---------------------------------------------------------------------------
// Bicubic table (src/table.cpp:42-158)
// ---------------------------------------------------------------------------
BT_DEV double spline1(double x, double v0, double v1, double d0, double d1)
{
    double a = 2*(v0-v1) + d0 + d1;
    double b = 3*(v1-v0) - 2*d0 - d1;
    double c = d0;
    double d = v0;
    return d + x*(c + x*(b + x*a));
}

BT_DEV double grad1(double x, double v0, double v1, double d0, double d1)
{
    double a = 2*(v0-v1) + d0 + d1;
    double b = 3*(v1-v0) - 2*d0 - d1;
    double c = d0;
    return c + x*(2*b + x*3*a);
}

// int(std::floor(u)) as baseline x86-64 evaluates it: NaN and values outside
// the int range become INT_MIN and fail the cell range test.  (The GPU's own
// conversion would map NaN to 0, a valid cell.)
BT_DEV int cell_index(double u)
{
    double f = floor(u);
    if (!(f >= -2147483648.0 && f < 2147483648.0)) return (int)0x80000000;
    return (int)f;
}

BT_DEV bool table_cell(const bt_surface* s, double x, double y, int& ix, int& iy)
{
    ix = cell_index((x - s->p[0])/s->p[2]);
    iy = cell_index((y - s->p[1])/s->p[3]);
    return !(ix >= s->n1-1 || ix < 0 || iy >= s->n2-1 || iy < 0);
}

BT_DEV double table_eval(const bt_surface* s, double x, double y)
{
    int ix, iy;
    if (!table_cell(s, x, y, ix, iy))
        return s->flags ? __builtin_nan("") : 0.0;
    const double dx = s->p[2], dy = s->p[3];
    const double* __restrict__ z = s->a0;
    const double* __restrict__ zx = s->a1;
    const double* __restrict__ zy = s->a2;
    const double* __restrict__ zxy = s->a3;
    double xgrid = s->p[0] + ix*dx;
    double ygrid = s->p[1] + iy*dy;
    double xfrac = (x - xgrid)/dx;
    double yfrac = (y - ygrid)/dy;
    const int64_t k0 = (int64_t)s->n1*iy + ix, k1 = k0 + s->n1;
    double val0 = spline1(xfrac, z[k0], z[k0+1], zx[k0]*dx, zx[k0+1]*dx);
    double val1 = spline1(xfrac, z[k1], z[k1+1], zx[k1]*dx, zx[k1+1]*dx);
    double der0 = spline1(xfrac, zy[k0], zy[k0+1], zxy[k0]*dx, zxy[k0+1]*dx);
    double der1 = spline1(xfrac, zy[k1], zy[k1+1], zxy[k1]*dx, zxy[k1+1]*dx);
    return spline1(yfrac, val0, val1, der0*dy, der1*dy);
}

BT_DEV void table_grad(const bt_surface* s, double x, double y,
                       double& dzdx, double& dzdy)
{
    int ix, iy;
    if (!table_cell(s, x, y, ix, iy)) {
        double v = s->flags ? __builtin_nan("") : 0.0;
        dzdx = v; dzdy = v;
        return;
    }
    const double dx = s->p[2], dy = s->p[3];
    const double* __restrict__ z = s->a0;
    const double* __restrict__ zx = s->a1;
    const double* __restrict__ zy = s->a2;
    const double* __restrict__ zxy = s->a3;
    double xgrid = s->p[0] + ix*dx;
    double ygrid = s->p[1] + iy*dy;
    double xfrac = (x - xgrid)/dx;
    double yfrac = (y - ygrid)/dy;
    const int64_t k0 = (int64_t)s->n1*iy + ix, k1 = k0 + s->n1;
    // the 16 table values of the cell, each read once
    const double z00 = z[k0], z01 = z[k0+1], z10 = z[k1], z11 = z[k1+1];
    const double x00 = zx[k0], x01 = zx[k0+1], x10 = zx[k1], x11 = zx[k1+1];
    const double y00 = zy[k0], y01 = zy[k0+1], y10 = zy[k1], y11 = zy[k1+1];
    const double c00 = zxy[k0], c01 = zxy[k0+1], c10 = zxy[k1], c11 = zxy[k1+1];

    double val0 = grad1(xfrac, z00, z01, x00*dx, x01*dx);
    double val1 = grad1(xfrac, z10, z11, x10*dx, x11*dx);
    double der0 = grad1(xfrac, y00, y01, c00*dx, c01*dx);
    double der1 = grad1(xfrac, y10, y11, c10*dx, c11*dx);
    dzdx = spline1(yfrac, val0, val1, der0*dy, der1*dy)/dx;

    val0 = grad1(yfrac, z00, z10, y00*dy, y10*dy);
    val1 = grad1(yfrac, z01, z11, y01*dy, y11*dy);
    der0 = grad1(yfrac, x00, x10, c00*dy, c10*dy);
    der1 = grad1(yfrac, x01, x11, c01*dy, c11*dy);
    dzdy = spline1(xfrac, val0, val1, der0*dx, der1*dx)/dy;
}

// ---------------------------------------------------------------------------
// 2-D Horner (src/polynomialSurface.cpp:50-66)
// ---------------------------------------------------------------------------
BT_DEV double horner2d(double x, double y, const double* __restrict__ c,
                       int nx, int ny)
{
    double r = 0.0;
    for (int i = ny-1; i >= 0; i--) {
        const double* row = c + (int64_t)i*nx;
        double h = 0.0;
        for (int j = nx-1; j >= 0; j--) { h *= y; h += row[j]; }
        r *= x;
        r += h;
    }
    return r;
}

// horner2d over trimmed bounds.  `trim` holds the last row with a
// coefficient other than +0.0 (-1 if none), then each row's length up to
// its last such coefficient.  The full loop starts each row (and the outer
// sum) from 0.0 and, over a run of +0.0 coefficients with a finite argument,
// stays exactly +0.0: (+-0)*y + (+0) == +0.  Starting after such a run from
// 0.0 therefore gives the same bits, and rows of only +0.0 add exactly +0.0
// where the full loop does.  A wave with any non-finite argument (0*inf is
// NaN inside the skipped steps) runs the full loop.
BT_DEV double horner2d_trim(double x, double y, const double* __restrict__ c,
                            int nx, int ny, const int* __restrict__ trim)
{
    if (!trim || !__all(__builtin_isfinite(x) && __builtin_isfinite(y)))
        return horner2d(x, y, c, nx, ny);
    double r = 0.0;
    for (int i = trim[0]; i >= 0; i--) {
        const double* row = c + (int64_t)i*nx;
        double h = 0.0;
        for (int j = trim[1 + i] - 1; j >= 0; j--) { h *= y; h += row[j]; }
        r *= x;
        r += h;
    }
    return r;
}

// Trim tables of a POLYNOMIAL record's coefficient, x-gradient and
// y-gradient tables, in that order, in the record's optional int32 table a3.
BT_DEV const int* poly_trim(const bt_surface* s, int which)
{
    const int* t = reinterpret_cast<const int*>(s->a3);
    if (!t) return nullptr;
    if (which == 0) return t;
    return t + (1 + s->n2) + (which - 1) * s->n2;
}

// table_eval and table_grad at one point, sharing the cell and its 16 table
// values; each result is bit-identical to the separate function's.
BT_DEV void table_eval_grad(const bt_surface* s, double x, double y, double& val,
                            double& dzdx, double& dzdy)
{
    int ix, iy;
    if (!table_cell(s, x, y, ix, iy)) {
        double v = s->flags ? __builtin_nan("") : 0.0;
        val = v; dzdx = v; dzdy = v;
        return;
    }
    const double dx = s->p[2], dy = s->p[3];
    const double* __restrict__ z = s->a0;
    const double* __restrict__ zx = s->a1;
    const double* __restrict__ zy = s->a2;
    const double* __restrict__ zxy = s->a3;
    double xgrid = s->p[0] + ix*dx;
    double ygrid = s->p[1] + iy*dy;
    double xfrac = (x - xgrid)/dx;
    double yfrac = (y - ygrid)/dy;
    const int64_t k0 = (int64_t)s->n1*iy + ix, k1 = k0 + s->n1;
    const double z00 = z[k0], z01 = z[k0+1], z10 = z[k1], z11 = z[k1+1];
    const double x00 = zx[k0], x01 = zx[k0+1], x10 = zx[k1], x11 = zx[k1+1];
    const double y00 = zy[k0], y01 = zy[k0+1], y10 = zy[k1], y11 = zy[k1+1];
    const double c00 = zxy[k0], c01 = zxy[k0+1], c10 = zxy[k1], c11 = zxy[k1+1];

    double val0 = spline1(xfrac, z00, z01, x00*dx, x01*dx);
    double val1 = spline1(xfrac, z10, z11, x10*dx, x11*dx);
    double der0 = spline1(xfrac, y00, y01, c00*dx, c01*dx);
    double der1 = spline1(xfrac, y10, y11, c10*dx, c11*dx);
    val = spline1(yfrac, val0, val1, der0*dy, der1*dy);

    val0 = grad1(xfrac, z00, z01, x00*dx, x01*dx);
    val1 = grad1(xfrac, z10, z11, x10*dx, x11*dx);
    der0 = grad1(xfrac, y00, y01, c00*dx, c01*dx);
    der1 = grad1(xfrac, y10, y11, c10*dx, c11*dx);
    dzdx = spline1(yfrac, val0, val1, der0*dy, der1*dy)/dx;

    val0 = grad1(yfrac, z00, z10, y00*dy, y10*dy);
    val1 = grad1(yfrac, z01, z11, y01*dy, y11*dy);
    der0 = grad1(yfrac, x00, x10, c00*dy, c10*dy);
    der1 = grad1(yfrac, x01, x11, c01*dy, c11*dy);
    dzdy = spline1(xfrac, val0, val1, der0*dx, der1*dx)/dy;
}

// ---------------------------------------------------------------------------
// Quadric family (src/quadric.cpp, src/asphere.cpp)
// ---------------------------------------------------------------------------
BT_DEV double quadric_sag(const bt_surface* s, double x, double y)
{
    const double R = s->p[0];
    double r2 = x*x + y*y;
    if (R != 0)
        return r2/(R*(1.+cr_sqrt(1.-r2*s->p[8])));
    return 0.0;
}

BT_DEV double quadric_dzdr(const bt_surface* s, double r)
{
    const double R = s->p[0];
    if (R != 0.0)
        return r/(R*cr_sqrt(1-r*r*s->p[8]));
    return 0.0;
}

BT_DEV double asphere_dzdr(const bt_surface* s, double r)
{
    double result = quadric_dzdr(s, r);
    double rr = r*r;
    double rrr = r;
    const double* __restrict__ dc = s->a1;
    const int n = s->n1;
    for (int i = 0; i < n; i++) {
        result += dc[i]*rrr;
        rrr *= rr;
    }
    return result;
}

BT_DEV double asphere_sag(const bt_surface* s, double x, double y)
{                                                    // asphere.cpp:37-46
    double r2 = x*x + y*y;
    double rr = 1.0;
    double result = quadric_sag(s, x, y);
    const double* __restrict__ c = s->a0;
    const int n = s->n1;
    for (int i = 0; i < n; i++) {
        rr *= r2;
        result += c[i]*rr;
    }
    return result;
}

// Normal of a Quadric (kAsphere false) or an Asphere (true): quadric.cpp:27-39,
// asphere.cpp:48-62.
template <bool kAsphere>
BT_DEV void conic_normal(const bt_surface* s, double x, double y,
                         double& nx, double& ny, double& nz)
{
    double r = cr_sqrt(x*x + y*y);
    if (r == 0.0) {
        nx = 0.0; ny = 0.0; nz = 1.0;
    } else {
        double dzdr = kAsphere ? asphere_dzdr(s, r) : quadric_dzdr(s, r);
        nz = cr_rsqrt(1+dzdr*dzdr);
        nx = -x/r*dzdr*nz;
        ny = -y/r*dzdr*nz;
    }
}

BT_DEV double poly_sag(const bt_surface* s, double x, double y)
{                                                    // polynomialSurface.cpp:35-37
    return horner2d_trim(x-s->p[0], y-s->p[1], s->a0, s->n1, s->n2, poly_trim(s, 0));
}

BT_DEV void poly_normal(const bt_surface* s, double x, double y,
                        double& nx, double& ny, double& nz)
{                                                    // polynomialSurface.cpp:39-48
    double gx = -horner2d_trim(x-s->p[0], y-s->p[1], s->a1, s->n1-1, s->n2-1, poly_trim(s, 1));
    double gy = -horner2d_trim(x-s->p[0], y-s->p[1], s->a2, s->n1-1, s->n2-1, poly_trim(s, 2));
    nz = cr_rsqrt(gx*gx + gy*gy + 1);
    nx = gx*nz;
    ny = gy*nz;
}

BT_DEV void table_normal(const bt_surface* s, double x, double y,
                         double& nx, double& ny, double& nz)
{                                                    // bicubic.cpp:23-40
    double dxdz, dydz;
    table_grad(s, x, y, dxdz, dydz);
    if (__builtin_isnan(dxdz)) {
        nx = __builtin_nan(""); ny = nx; nz = nx;
        return;
    }
    nz = cr_rsqrt(1 + dxdz*dxdz + dydz*dydz);
    nx = -dxdz*nz;
    ny = -dydz*nz;
}

// Bicubic sag and normal at one point, sharing the cell (bit-identical to
// table_eval and table_normal).
BT_DEV void table_sag_normal(const bt_surface* s, double x, double y, double& sag,
                             double& nx, double& ny, double& nz)
{                                                    // bicubic.cpp:19-40
    double dxdz, dydz;
    table_eval_grad(s, x, y, sag, dxdz, dydz);
    if (__builtin_isnan(dxdz)) {
        nx = __builtin_nan(""); ny = nx; nz = nx;
        return;
    }
    nz = cr_rsqrt(1 + dxdz*dxdz + dydz*dydz);
    nx = -dxdz*nz;
    ny = -dydz*nz;
}

// ---------------------------------------------------------------------------
// Surface kinds a kernel is compiled for.
//  * kKindsAll: every kind, Sums of any children (a switch over the kind of
//    each child inside the Sum's loop);
//  * kKindsAnalytic: Plane, Sphere, Paraboloid, Quadric, Asphere and Tilted,
//    no Sums -- the Bicubic, polynomial and Sum code paths are left out,
//    which lets a plan kernel for such optics keep far fewer registers;
//  * kKindsSumAPB: those analytic leaves plus Sums of exactly (Asphere,
//    Polynomial, Bicubic), in that order (a perturbed conic mirror: figure,
//    low-order Zernike error, measured high-frequency map).  The Sum's three
//    children are evaluated by their own functions with no kind switch, so
//    each Newton step carries only the registers those three need;
//  * kKindsFlatSums: every kind, Sums whose children are all leaves (the
//    evaluation of nested Sums is left out: it costs the one-level Sum path
//    registers).
// kKindsAll also evaluates Sums of Sums, to kMaxSumDepth levels below a Sum.
// Under a restricted set any other kind (or Sum layout) evaluates to NaN and
// so fails.
// ---------------------------------------------------------------------------
constexpr int kKindsAll = 0;
constexpr int kKindsAnalytic = 1;
constexpr int kKindsSumAPB = 2;
constexpr int kKindsFlatSums = 3;

// The kind set of leaves (records that are not Sums) under set K.
constexpr int leaf_set(int K)
{
    return (K == kKindsAll || K == kKindsFlatSums) ? kKindsAll : kKindsAnalytic;
}

// A Sum record laid out as kKindsSumAPB requires (records s+1..s+3 follow it).
BT_DEV bool sum_is_apb(const bt_surface* s)
{
    return s->nchild == 3 && s[1].kind == BT_ASPHERE && s[2].kind == BT_POLYNOMIAL &&
           s[3].kind == BT_BICUBIC;
}

// ---------------------------------------------------------------------------
// Leaf surfaces: sag and normal
// ---------------------------------------------------------------------------
template <int K = kKindsAll>
BT_DEV double leaf_sag(const bt_surface* s, double x, double y)
{
    switch (s->kind) {
    case BT_PLANE:                                   // plane.cpp:15-17
        return 0.0;
    case BT_SPHERE:                                  // sphere.cpp:17-21
        if (s->p[0] != 0)
            return s->p[0]*(1-cr_sqrt(1-(x*x + y*y)*s->p[3]));
        return 0.0;
    case BT_PARABOLOID:                              // paraboloid.cpp:17-23
        if (s->p[0] != 0) {
            double r2 = x*x + y*y;
            return r2*s->p[2];
        }
        return 0.0;
    case BT_QUADRIC:                                 // quadric.cpp:18-25
        return quadric_sag(s, x, y);
    case BT_ASPHERE:                                 // asphere.cpp:37-46
        return asphere_sag(s, x, y);
    case BT_TILTED:                                  // tilted.cpp:16-18
        return x*s->p[0] + y*s->p[1];
    case BT_BICUBIC:                                 // bicubic.cpp:19-21
        if constexpr (K == kKindsAll) return table_eval(s, x, y);
        break;
    case BT_POLYNOMIAL:                              // polynomialSurface.cpp:35-37
        if constexpr (K == kKindsAll) return poly_sag(s, x, y);
        break;
    }
    return __builtin_nan("");
}

template <int K = kKindsAll>
BT_DEV void leaf_normal(const bt_surface* s, double x, double y,
                        double& nx, double& ny, double& nz)
{
    switch (s->kind) {
    case BT_PLANE:                                   // plane.cpp:19-23
        nx = 0.0; ny = 0.0; nz = 1.0;
        return;
    case BT_SPHERE: {                                // sphere.cpp:23-35, 66-69
        double rsqr = x*x + y*y;
        if (s->p[0] == 0.0 || rsqr == 0.0) {
            nx = 0.0; ny = 0.0; nz = 1.0;
        } else {
            double r = cr_sqrt(rsqr);
            double rat = r*s->p[2];
            double dzdr = rat/cr_sqrt(1-rat*rat);
            nz = cr_rsqrt(1+dzdr*dzdr);
            nx = -x/r*dzdr*nz;
            ny = -y/r*dzdr*nz;
        }
        return;
    }
    case BT_PARABOLOID: {                            // paraboloid.cpp:25-37
        if (s->p[0] == 0) {
            nx = 0.0; ny = 0.0; nz = 1.0;
        } else {
            const double Rinv = s->p[1];
            nz = cr_rsqrt(1+(x*x+y*y)*Rinv*Rinv);
            nx = -x*Rinv*nz;
            ny = -y*Rinv*nz;
        }
        return;
    }
    case BT_QUADRIC:                                 // quadric.cpp:27-39
    case BT_ASPHERE: {                               // asphere.cpp:48-62
        double r = cr_sqrt(x*x + y*y);
        if (r == 0.0) {
            nx = 0.0; ny = 0.0; nz = 1.0;
        } else {
            double dzdr = (s->kind == BT_QUADRIC) ? quadric_dzdr(s, r)
                                                  : asphere_dzdr(s, r);
            nz = cr_rsqrt(1+dzdr*dzdr);
            nx = -x/r*dzdr*nz;
            ny = -y/r*dzdr*nz;
        }
        return;
    }
    case BT_TILTED: {                                // tilted.cpp:20-27
        const double tx = s->p[0], ty = s->p[1];
        nx = -tx;
        ny = -ty;
        nz = cr_sqrt(1 - tx*tx - ty*ty);
        return;
    }
    case BT_BICUBIC:                                 // bicubic.cpp:23-40
        if constexpr (K == kKindsAll) { table_normal(s, x, y, nx, ny, nz); return; }
        break;
    case BT_POLYNOMIAL:                              // polynomialSurface.cpp:39-48
        if constexpr (K == kKindsAll) { poly_normal(s, x, y, nx, ny, nz); return; }
        break;
    }
    nx = __builtin_nan(""); ny = nx; nz = nx;
}

// Sag and normal at one point (the reference evaluates them separately; the
// results here are bit-identical, with shared work evaluated once).
template <int K = kKindsAll>
BT_DEV void leaf_sag_normal(const bt_surface* s, double x, double y, double& sag,
                            double& nx, double& ny, double& nz)
{
    if (K == kKindsAll && s->kind == BT_BICUBIC) {
        table_sag_normal(s, x, y, sag, nx, ny, nz);
        return;
    }
    sag = leaf_sag<K>(s, x, y);
    leaf_normal<K>(s, x, y, nx, ny, nz);
}

BT_DEV void add_slope(double& sx, double& sy, double nx, double ny, double nz)
{
    sx += nx/nz;
    sy += ny/nz;
}

BT_DEV void sum_normal_from_slopes(double sx, double sy, double& nx, double& ny, double& nz)
{
    nz = cr_rsqrt(sx*sx + sy*sy + 1);
    nx = sx*nz;
    ny = sy*nz;
}

// Sums of any kinds (kKindsAll).  A Sum record is followed by its children,
// each a leaf record or a Sum with its own subtree; a Sum's n1 counts the
// records below it, so the next sibling starts 1 + n1 records on.  Children
// that are Sums are evaluated by the same rules (sum.cpp:23-45 calls the
// child's own sag / normal), to kMaxSumDepth levels of nesting below a Sum
// (the host refuses deeper trees; a deeper child evaluates to NaN).
constexpr int kMaxSumDepth = 3;

BT_DEV const bt_surface* next_child(const bt_surface* c)
{
    return c + (c->kind == BT_SUM ? 1 + c->n1 : 1);
}

template <int D>
BT_DEV double sum_sag_all(const bt_surface* s, double x, double y)
{                                                    // sum.cpp:23-28
    double result = 0.0;
    const bt_surface* c = s + 1;
    const int nc = s->nchild;
    for (int i = 0; i < nc; i++, c = next_child(c)) {
        if (c->kind != BT_SUM) {
            result += leaf_sag<kKindsAll>(c, x, y);
        } else {
            if constexpr (D > 0) result += sum_sag_all<D - 1>(c, x, y);
            else result += __builtin_nan("");
        }
    }
    return result;
}

template <int D>
BT_DEV void sum_normal_all(const bt_surface* s, double x, double y,
                           double& nx, double& ny, double& nz)
{                                                    // sum.cpp:30-45
    double sx = 0.0, sy = 0.0, tnx, tny, tnz;
    const bt_surface* c = s + 1;
    const int nc = s->nchild;
    for (int i = 0; i < nc; i++, c = next_child(c)) {
        if (c->kind != BT_SUM) {
            leaf_normal<kKindsAll>(c, x, y, tnx, tny, tnz);
        } else {
            if constexpr (D > 0) sum_normal_all<D - 1>(c, x, y, tnx, tny, tnz);
            else { tnx = __builtin_nan(""); tny = tnx; tnz = tnx; }
        }
        add_slope(sx, sy, tnx, tny, tnz);
    }
    sum_normal_from_slopes(sx, sy, nx, ny, nz);
}

template <int D>
BT_DEV void sum_sag_normal_all(const bt_surface* s, double x, double y, double& sag,
                               double& nx, double& ny, double& nz)
{
    double result = 0.0, sx = 0.0, sy = 0.0, ts, tnx, tny, tnz;
    const bt_surface* c = s + 1;
    const int nc = s->nchild;
    for (int i = 0; i < nc; i++, c = next_child(c)) {
        if (c->kind != BT_SUM) {
            leaf_sag_normal<kKindsAll>(c, x, y, ts, tnx, tny, tnz);
        } else {
            if constexpr (D > 0) sum_sag_normal_all<D - 1>(c, x, y, ts, tnx, tny, tnz);
            else { ts = __builtin_nan(""); tnx = ts; tny = ts; tnz = ts; }
        }
        result += ts;
        add_slope(sx, sy, tnx, tny, tnz);
    }
    sag = result;
    sum_normal_from_slopes(sx, sy, nx, ny, nz);
}

// Sum records are followed by their children (sum.cpp:23-45).  The Sum's
// sag is 0 + child sags in child order, its normal from the sum of the
// children's slopes (-n_x/n_z, -n_y/n_z) in child order, so the unrolled
// kKindsSumAPB forms below are the same operations as the loops.
template <int K = kKindsAll>
BT_DEV double surf_sag(const bt_surface* s, double x, double y)
{
    constexpr int L = leaf_set(K);
    if (K == kKindsAnalytic || s->kind != BT_SUM) return leaf_sag<L>(s, x, y);
    double result = 0.0;
    if constexpr (K == kKindsSumAPB) {
        if (!sum_is_apb(s)) return __builtin_nan("");
        result += asphere_sag(s + 1, x, y);
        result += poly_sag(s + 2, x, y);
        result += table_eval(s + 3, x, y);
        return result;
    } else if constexpr (K == kKindsFlatSums) {
        const int nc = s->nchild;
        for (int i = 0; i < nc; i++)
            result += leaf_sag<L>(s + 1 + i, x, y);
        return result;
    } else {
        return sum_sag_all<kMaxSumDepth>(s, x, y);
    }
}

template <int K = kKindsAll>
BT_DEV void surf_normal(const bt_surface* s, double x, double y,
                        double& nx, double& ny, double& nz)
{
    constexpr int L = leaf_set(K);
    if (K == kKindsAnalytic || s->kind != BT_SUM) { leaf_normal<L>(s, x, y, nx, ny, nz); return; }
    double sx = 0.0, sy = 0.0;
    double tnx, tny, tnz;
    if constexpr (K == kKindsSumAPB) {
        if (!sum_is_apb(s)) { nx = __builtin_nan(""); ny = nx; nz = nx; return; }
        conic_normal<true>(s + 1, x, y, tnx, tny, tnz);
        add_slope(sx, sy, tnx, tny, tnz);
        poly_normal(s + 2, x, y, tnx, tny, tnz);
        add_slope(sx, sy, tnx, tny, tnz);
        table_normal(s + 3, x, y, tnx, tny, tnz);
        add_slope(sx, sy, tnx, tny, tnz);
        sum_normal_from_slopes(sx, sy, nx, ny, nz);
        return;
    } else if constexpr (K == kKindsFlatSums) {
        const int nc = s->nchild;
        for (int i = 0; i < nc; i++) {
            leaf_normal<L>(s + 1 + i, x, y, tnx, tny, tnz);
            add_slope(sx, sy, tnx, tny, tnz);
        }
        sum_normal_from_slopes(sx, sy, nx, ny, nz);
    } else {
        sum_normal_all<kMaxSumDepth>(s, x, y, nx, ny, nz);
    }
}

template <int K = kKindsAll>
BT_DEV void surf_sag_normal(const bt_surface* s, double x, double y, double& sag,
                            double& nx, double& ny, double& nz)
{
    constexpr int L = leaf_set(K);
    if (K == kKindsAnalytic || s->kind != BT_SUM) {
        leaf_sag_normal<L>(s, x, y, sag, nx, ny, nz);
        return;
    }
    double result = 0.0, sx = 0.0, sy = 0.0;
    double ts, tnx, tny, tnz;
    if constexpr (K == kKindsSumAPB) {
        if (!sum_is_apb(s)) { sag = __builtin_nan(""); nx = sag; ny = sag; nz = sag; return; }
        result += asphere_sag(s + 1, x, y);
        conic_normal<true>(s + 1, x, y, tnx, tny, tnz);
        add_slope(sx, sy, tnx, tny, tnz);
        result += poly_sag(s + 2, x, y);
        poly_normal(s + 2, x, y, tnx, tny, tnz);
        add_slope(sx, sy, tnx, tny, tnz);
        table_sag_normal(s + 3, x, y, ts, tnx, tny, tnz);
        result += ts;
        add_slope(sx, sy, tnx, tny, tnz);
        sag = result;
        sum_normal_from_slopes(sx, sy, nx, ny, nz);
        return;
    } else if constexpr (K == kKindsFlatSums) {
        const int nc = s->nchild;
        for (int i = 0; i < nc; i++) {
            leaf_sag_normal<L>(s + 1 + i, x, y, ts, tnx, tny, tnz);
            result += ts;
            add_slope(sx, sy, tnx, tny, tnz);
        }
        sag = result;
        sum_normal_from_slopes(sx, sy, nx, ny, nz);
    } else {
        sum_sag_normal_all<kMaxSumDepth>(s, x, y, sag, nx, ny, nz);
    }
}

// Surface evaluators for the Newton loops.  SurfEval<K> dispatches on the
// record's kind at every evaluation; where the caller already knows the kind
// (an Asphere leaf, a Sum laid out as kKindsSumAPB requires) the loop is
// instantiated with that kind's functions directly, so it carries no switch
// over the other kinds.  Every evaluator computes exactly what surf_sag /
// surf_normal / surf_sag_normal compute for that record.
template <int K>
struct SurfEval {
    static BT_DEV double sag(const bt_surface* s, double x, double y)
    {
        return surf_sag<K>(s, x, y);
    }
    static BT_DEV void normal(const bt_surface* s, double x, double y, double& nx, double& ny,
                              double& nz)
    {
        surf_normal<K>(s, x, y, nx, ny, nz);
    }
    static BT_DEV void sag_normal(const bt_surface* s, double x, double y, double& sag,
                                  double& nx, double& ny, double& nz)
    {
        surf_sag_normal<K>(s, x, y, sag, nx, ny, nz);
    }
};

// A record known to be a leaf (not a Sum).
template <int K>
struct LeafEval {
    static BT_DEV double sag(const bt_surface* s, double x, double y)
    {
        return leaf_sag<K>(s, x, y);
    }
    static BT_DEV void normal(const bt_surface* s, double x, double y, double& nx, double& ny,
                              double& nz)
    {
        leaf_normal<K>(s, x, y, nx, ny, nz);
    }
    static BT_DEV void sag_normal(const bt_surface* s, double x, double y, double& sag,
                                  double& nx, double& ny, double& nz)
    {
        leaf_sag_normal<K>(s, x, y, sag, nx, ny, nz);
    }
};

struct AsphereEval {                                 // asphere.cpp:37-62
    static BT_DEV double sag(const bt_surface* s, double x, double y)
    {
        return asphere_sag(s, x, y);
    }
    static BT_DEV void normal(const bt_surface* s, double x, double y, double& nx, double& ny,
                              double& nz)
    {
        conic_normal<true>(s, x, y, nx, ny, nz);
    }
    static BT_DEV void sag_normal(const bt_surface* s, double x, double y, double& sag,
                                  double& nx, double& ny, double& nz)
    {
        sag = asphere_sag(s, x, y);
        conic_normal<true>(s, x, y, nx, ny, nz);
    }
};

// A Sum whose layout sum_is_apb has accepted (sum.cpp:23-45).
struct SumAPBEval {
    static BT_DEV double sag(const bt_surface* s, double x, double y)
    {
        double result = 0.0;
        result += asphere_sag(s + 1, x, y);
        result += poly_sag(s + 2, x, y);
        result += table_eval(s + 3, x, y);
        return result;
    }
    static BT_DEV void normal(const bt_surface* s, double x, double y, double& nx, double& ny,
                              double& nz)
    {
        double sx = 0.0, sy = 0.0, tnx, tny, tnz;
        conic_normal<true>(s + 1, x, y, tnx, tny, tnz);
        add_slope(sx, sy, tnx, tny, tnz);
        poly_normal(s + 2, x, y, tnx, tny, tnz);
        add_slope(sx, sy, tnx, tny, tnz);
        table_normal(s + 3, x, y, tnx, tny, tnz);
        add_slope(sx, sy, tnx, tny, tnz);
        sum_normal_from_slopes(sx, sy, nx, ny, nz);
    }
    static BT_DEV void sag_normal(const bt_surface* s, double x, double y, double& sag,
                                  double& nx, double& ny, double& nz)
    {
        double result = 0.0, sx = 0.0, sy = 0.0, ts, tnx, tny, tnz;
        result += asphere_sag(s + 1, x, y);
        conic_normal<true>(s + 1, x, y, tnx, tny, tnz);
        add_slope(sx, sy, tnx, tny, tnz);
        result += poly_sag(s + 2, x, y);
        poly_normal(s + 2, x, y, tnx, tny, tnz);
        add_slope(sx, sy, tnx, tny, tnz);
        table_sag_normal(s + 3, x, y, ts, tnx, tny, tnz);
        result += ts;
        add_slope(sx, sy, tnx, tny, tnz);
        sag = result;
        sum_normal_from_slopes(sx, sy, nx, ny, nz);
    }
};

BT_DEV bool same_bits(double a, double b)
{
    return __double_as_longlong(a) == __double_as_longlong(b);
}

// Tangent-plane Newton with exactly `niter` iterations (surface.cpp:22-49).
//
// Each iteration is a pure function of dt (the ray origin and direction are
// fixed): dt -> rP(dt), sag(rP), normal(rP) -> dt'.  So once an iterate
// repeats bit for bit -- a fixed point dt' == dt, or a two-cycle between
// neighbouring doubles dt'' == dt -- every remaining iterate is known, and the
// loop stops with exactly the state (dt, rP, sag) that running all `niter`
// iterations would leave.  Results are bit-identical to the full loop.
// The last surface normal Newton evaluated and the dt of the point it was
// evaluated at: a reflection or refraction at that same point reuses it
// (surf_normal of bit-identical arguments is bit-identical).
struct NormalMemo {
    double dt, nx, ny, nz;
    bool ok;
};

template <class Ev>
BT_DEV bool newton(const bt_surface* s, double x, double y, double z,
                   double vx, double vy, double vz, double& dt, int niter,
                   NormalMemo* memo = nullptr)
{
    double rPx = x+vx*dt;
    double rPy = y+vy*dt;
    double rPz = z+vz*dt;
    double sz = Ev::sag(s, rPx, rPy);
    double pdt = 0.0;            // the iterate before the current one (iter >= 1)
    for (int iter = 0; iter < niter; iter++) {
        double nx, ny, nz;
        Ev::normal(s, rPx, rPy, nx, ny, nz);
        if (memo) *memo = NormalMemo{dt, nx, ny, nz, true};
        double dtn = (rPx-x)*nx + (rPy-y)*ny + (sz-z)*nz;
        dtn /= (nx*vx + ny*vy + nz*vz);
        if (same_bits(dtn, dt)) break;            // fixed point: state unchanged
        const double odt = dt, oz = rPz, osz = sz;
        dt = dtn;
        rPx = x+vx*dt;
        rPy = y+vy*dt;
        rPz = z+vz*dt;
        sz = Ev::sag(s, rPx, rPy);
        if (iter >= 1 && same_bits(dt, pdt)) {
            // two-cycle: iterates alternate between this state and the
            // previous one; after the remaining (niter-1-iter) iterations the
            // loop ends on this state if that count is even, else on the other
            if ((niter - 1 - iter) & 1) { dt = odt; rPz = oz; sz = osz; }
            break;
        }
        pdt = odt;
    }
    return fabs(sz-rPz) < 1e-12;
}

// std::max(a, b) = (a < b) ? b : a; std::min(a, b) = (b < a) ? b : a.
BT_DEV double pick_root(double vzR, double dt1, double dt2)
{
    if (vzR < 0) return (dt1 < dt2) ? dt2 : dt1;
    return (dt2 < dt1) ? dt2 : dt1;
}

BT_DEV bool quadric_tti(const bt_surface* s, double x, double y, double z,
                        double vx, double vy, double vz, double& dt)
{                                                    // quadric.cpp:41-70
    double z0term = z-s->p[6];
    double vrr0 = vx*x + vy*y;
    double a = vz*vz + (vx*vx+vy*vy)*s->p[5];
    double b = 2*(vz*z0term + vrr0*s->p[5]);
    double c = z0term*z0term - s->p[7] + (x*x + y*y)*s->p[5];
    double discriminant = b*b - 4*a*c;
    if (discriminant < 0)
        return false;
    double dt1;
    if (b > 0)
        dt1 = (-b - cr_sqrt(discriminant)) / (2*a);
    else
        dt1 = 2*c / (-b + cr_sqrt(discriminant));
    double dt2 = c / (a*dt1);
    double z1 = z + vz*dt1;
    double z2 = z + vz*dt2;
    dt = (fabs(z1) < fabs(z2)) ? dt1 : dt2;
    return true;
}

template <int K = kKindsAll>
BT_DEV bool leaf_tti(const bt_surface* s, double x, double y, double z,
                     double vx, double vy, double vz, double& dt, int niter,
                     NormalMemo* memo = nullptr)
{
    switch (s->kind) {
    case BT_PLANE:                                   // plane.cpp:25-31
        if (vz == 0) return false;
        dt = -z/vz;
        return true;
    case BT_SPHERE: {                                // sphere.cpp:37-64
        const double R = s->p[0];
        double vr2 = vx*vx + vy*vy;
        double vz2 = vz*vz;
        double vrr0 = vx*x + vy*y;
        double r02 = x*x + y*y;
        double z0term = z-R;
        double a = vz2 + vr2;
        double b = 2*vz*z0term + 2*vrr0;
        double c = z0term*z0term - s->p[1] + r02;
        double discriminant = b*b - 4*a*c;
        if (discriminant < 0) return false;
        double dt1;
        if (b > 0)
            dt1 = (-b - cr_sqrt(discriminant)) / (2*a);
        else
            dt1 = 2*c / (-b + cr_sqrt(discriminant));
        double dt2 = c / (a*dt1);
        dt = pick_root(vz*R, dt1, dt2);
        return true;
    }
    case BT_PARABOLOID: {                            // paraboloid.cpp:39-71
        double a = (vx*vx + vy*vy)*s->p[2];
        double b = (x*vx + y*vy)*s->p[1] - vz;
        double c = (x*x + y*y)*s->p[2] - z;
        if (a == 0) {
            dt = -c/b;
            return true;
        }
        double discriminant = b*b - 4*a*c;
        if (discriminant < 0) return false;
        double dt1;
        if (b > 0)
            dt1 = (-b - cr_sqrt(discriminant)) / (2*a);
        else
            dt1 = 2*c / (-b + cr_sqrt(discriminant));
        double dt2 = c / (a*dt1);
        dt = pick_root(vz*s->p[0], dt1, dt2);
        return true;
    }
    case BT_QUADRIC:
        return quadric_tti(s, x, y, z, vx, vy, vz, dt);
    case BT_ASPHERE:                                 // asphere.cpp:75-85
        if (!quadric_tti(s, x, y, z, vx, vy, vz, dt)) return false;
        return newton<AsphereEval>(s, x, y, z, vx, vy, vz, dt, niter, memo);
    case BT_TILTED: {                                // tilted.cpp:29-37
        double num = z - x*s->p[0] - y*s->p[1];
        double den = vx*s->p[0] + vy*s->p[1] - vz;
        if (den == 0) return false;
        dt = num/den;
        return true;
    }
    case BT_BICUBIC:
    case BT_POLYNOMIAL:                              // surface.cpp:22-49
        if constexpr (K == kKindsAll)
            return newton<LeafEval<K>>(s, x, y, z, vx, vy, vz, dt, niter, memo);
        break;
    }
    return false;
}

// sum.cpp:47-56: the first child's intersection (itself a Sum's, to the
// nesting depth kKindsAll allows) seeds Newton on the Sum.
template <int D>
BT_DEV bool sum_tti_all(const bt_surface* s, double x, double y, double z,
                        double vx, double vy, double vz, double& dt, int niter,
                        NormalMemo* memo)
{
    const bt_surface* c = s + 1;
    bool ok = false;
    if (c->kind != BT_SUM) {
        ok = leaf_tti<kKindsAll>(c, x, y, z, vx, vy, vz, dt, niter);
    } else {
        if constexpr (D > 0) ok = sum_tti_all<D - 1>(c, x, y, z, vx, vy, vz, dt, niter, nullptr);
    }
    if (!ok) return false;
    return newton<SurfEval<kKindsAll>>(s, x, y, z, vx, vy, vz, dt, niter, memo);
}

template <int K = kKindsAll>
BT_DEV bool surf_tti(const bt_surface* s, double x, double y, double z,
                     double vx, double vy, double vz, double& dt, int niter,
                     NormalMemo* memo = nullptr)
{
    constexpr int L = leaf_set(K);
    if (K == kKindsAnalytic || s->kind != BT_SUM)
        return leaf_tti<L>(s, x, y, z, vx, vy, vz, dt, niter, memo);
    if constexpr (K == kKindsSumAPB) {
        if (!sum_is_apb(s)) return false;
        if (!leaf_tti<L>(s + 1, x, y, z, vx, vy, vz, dt, niter)) return false;
        return newton<SumAPBEval>(s, x, y, z, vx, vy, vz, dt, niter, memo);
    } else if constexpr (K == kKindsFlatSums) {
        if (!leaf_tti<L>(s + 1, x, y, z, vx, vy, vz, dt, niter)) return false;
        return newton<SurfEval<K>>(s, x, y, z, vx, vy, vz, dt, niter, memo);
    } else {
        return sum_tti_all<kMaxSumDepth>(s, x, y, z, vx, vy, vz, dt, niter, memo);
    }
}

// The surface's analytic base start: the root of a Sum's first child or of an
// Asphere's conic (the starts of sum.cpp:53 and asphere.cpp:82, without the
// Newton refinement that follows them); surfaces without one start at 0.
template <int K = kKindsAll>
BT_DEV bool conic_start(const bt_surface* s, double x, double y, double z, double vx,
                        double vy, double vz, double& dt)
{
    const bt_surface* c = s;
    if (K != kKindsAnalytic)
        while (c->kind == BT_SUM) c = c + 1;     // the first leaf of nested Sums
    dt = 0.0;
    switch (c->kind) {
    case BT_ASPHERE:
    case BT_QUADRIC:
        return quadric_tti(c, x, y, z, vx, vy, vz, dt);
    case BT_PLANE:
    case BT_SPHERE:
    case BT_PARABOLOID:
    case BT_TILTED:
        return leaf_tti<leaf_set(K)>(c, x, y, z, vx, vy, vz, dt, 0);
    default:
        return true;
    }
}


BT_DEV bool newton_surface(const bt_surface* s)
{
    return s->kind == BT_SUM || s->kind == BT_ASPHERE || s->kind == BT_BICUBIC ||
           s->kind == BT_POLYNOMIAL;
}

constexpr double kFastTol = 1e-14;   // sag gap accepted by the fast starts [m]

// Newton on the full surface from `dt` until |sag - z| <= kFastTol, with at
// most `max_iter` updates (the reference's niter: a fast start never takes
// more iterations than the reference would, so a ray whose reference
// iteration from the same start has not converged by then is handed back to
// the reference path, which decides it).  At 1e-13 m the residual position
// errors, carried through a three-mirror optic, reach ~1e-14 in the final
// direction cosines; 1e-14 m keeps them an order of magnitude below that.
template <class Ev>
BT_DEV bool newton_fast(const bt_surface* s, double x, double y, double z,
                        double vx, double vy, double vz, double& dt, NormalMemo* memo,
                        int max_iter)
{
    double rPx = x+vx*dt, rPy = y+vy*dt, rPz = z+vz*dt;
    double sz, nx, ny, nz;
    Ev::sag_normal(s, rPx, rPy, sz, nx, ny, nz);
    for (int it = 0; it < max_iter; it++) {
        if (fabs(sz-rPz) <= kFastTol) break;
        double d = (rPx-x)*nx + (rPy-y)*ny + (sz-z)*nz;
        dt = d / (nx*vx + ny*vy + nz*vz);
        rPx = x+vx*dt;
        rPy = y+vy*dt;
        rPz = z+vz*dt;
        Ev::sag_normal(s, rPx, rPy, sz, nx, ny, nz);
    }
    if (!(fabs(sz-rPz) <= kFastTol)) return false;
    *memo = NormalMemo{dt, nx, ny, nz, true};
    return true;
}

// One Newton update of `dt` on an Asphere alone (its sag and normal at the
// ray's point at `dt`); `dt` is kept where the update is not finite.
BT_DEV void asphere_update(const bt_surface* c, double x, double y, double z, double vx,
                           double vy, double vz, double& dt)
{
    const double rPx = x+vx*dt, rPy = y+vy*dt;
    double sz, nx, ny, nz;
    AsphereEval::sag_normal(c, rPx, rPy, sz, nx, ny, nz);
    const double d = (rPx-x)*nx + (rPy-y)*ny + (sz-z)*nz;
    const double t = d / (nx*vx + ny*vy + nz*vz);
    if (t == t) dt = t;
}

// Intersection with a fast start, then Newton on the full surface to the
// fast tolerance.  Starts:
// * BT_START_CONIC on a Sum whose first child is an Asphere: that Asphere's
//   conic root moved by one Newton update on the Asphere (the Sum's other
//   terms then hold the whole remaining gap, so the Sum's own, costlier,
//   iteration takes one update where the conic root needs up to two);
// * BT_START_CONIC on another Sum: the reference's own start, its first
//   child's root (sum.cpp:53), so the fast iteration runs the reference's
//   iterates;
// * otherwise the base conic's root (the ray origin where the surface has
//   no conic base) plus `correction` (the network's prediction for
//   BT_START_MLP, 0 for BT_START_CONIC).
// The reference path decides every ray the fast start does not settle.
template <int K = kKindsAll>
BT_DEV bool fast_tti(const bt_surface* s, double x, double y, double z, double vx,
                     double vy, double vz, double& dt, int niter, double correction,
                     NormalMemo* memo)
{
    if (newton_surface(s)) {
        double d0 = 0.0;
        bool ok;
        if (K != kKindsAnalytic && correction == 0.0 && s->kind == BT_SUM) {
            const bt_surface* c = s + 1;
            if (c->kind == BT_ASPHERE) {
                ok = quadric_tti(c, x, y, z, vx, vy, vz, d0);
                if (ok) asphere_update(c, x, y, z, vx, vy, vz, d0);
            } else if constexpr (K != kKindsSumAPB) {
                ok = surf_tti<K>(c, x, y, z, vx, vy, vz, d0, niter);
            } else {
                ok = false;                     // not a Sum(Asphere, Polynomial, Bicubic)
            }
        } else {
            ok = conic_start<K>(s, x, y, z, vx, vy, vz, d0);
            d0 += correction;
        }
        bool conv = false;
        if (ok && d0 == d0) {
            if (s->kind == BT_ASPHERE)
                conv = newton_fast<AsphereEval>(s, x, y, z, vx, vy, vz, d0, memo, niter);
            else if constexpr (K == kKindsSumAPB)
                conv = sum_is_apb(s) &&
                       newton_fast<SumAPBEval>(s, x, y, z, vx, vy, vz, d0, memo, niter);
            else
                conv = newton_fast<SurfEval<K>>(s, x, y, z, vx, vy, vz, d0, memo, niter);
        }
        if (conv) {
            dt = d0;
            return true;
        }
    }
    dt = 0.0;
    return surf_tti<K>(s, x, y, z, vx, vy, vz, dt, niter, memo);
}

// ---------------------------------------------------------------------------
// Media (src/medium.cpp)
// ---------------------------------------------------------------------------
BT_DEV double medium_n(const bt_medium* m, double w)
{
    switch (m->kind) {
    case BT_MED_CONST:                               // medium.cpp:30-32
        return m->p[0];
    case BT_MED_TABLE: {                             // medium.cpp:89-105
        const double* __restrict__ args = m->a0;
        const double* __restrict__ vals = m->a1;
        const int size = m->n;
        if (w < args[0]) return __builtin_nan("");
        if (w > args[size-1]) return __builtin_nan("");
        int u;
        for (u = 1; u < size; u++)
            if (w < args[u]) break;
        double out = (w - args[u-1]);
        out *= (vals[u] - vals[u-1]);
        out /= (args[u] - args[u-1]);
        out += vals[u-1];
        return out;
    }
    case BT_MED_SELLMEIER: {                         // medium.cpp:145-149
        const double* p = m->p;
        double x = w*w*1e12;
        return cr_sqrt(1.0 + p[0]*x/(x-p[3]) + p[1]*x/(x-p[4]) + p[2]*x/(x-p[5]));
    }
    case BT_MED_SUMITA: {                            // medium.cpp:186-191
        const double* A = m->p;
        double x = w*w*1e12;
        double y = 1./x;
        return cr_sqrt(A[0] + A[1]*x + y*(A[2] + y*(A[3] + y*(A[4] + y*A[5]))));
    }
    case BT_MED_AIR: {                               // medium.cpp:237-245
        const double P = m->p[0], T = m->p[1], W = m->p[2];
        double s2 = 1e-12 / (w*w);
        double n_minus_one = (64.328 + (29498.1 / (146.0 - s2))
                              + (255.4 / (41.0 - s2))) * 1.e-6;
        n_minus_one *= P * (1.0 + (1.049 - 0.0157 * T) * 1.e-6 * P)
                       / (720.883 * (1.0 + 0.003661 * T));
        n_minus_one -= (0.0624 - 0.000680 * s2)/(1.0 + 0.003661 * T) * W * 1.e-6;
        return 1+n_minus_one;
    }
    }
    return __builtin_nan("");
}

// ---------------------------------------------------------------------------
// hypot as the reference gets it from x86-64 libm (glibc >= 2.35): the
// corrected square root of C. F. Borges, "An improved algorithm for
// hypot(a, b)" (2019), in its form without fused multiply-add, with
// power-of-two rescaling far from 1.  Bit-identical to libm's hypot,
// including the ~0.4 % of arguments where that is not the correctly
// rounded value -- which decides flags of rays that land exactly on a
// circular obscuration edge (a hexapolar ring at the pupil radius).
// Needs -ffp-contract=off (no fusing of the products below).
// ---------------------------------------------------------------------------
BT_DEV double hypot_kernel(double ax, double ay)    // ax >= ay >= 0
{
    double h = cr_sqrt(ax*ax + ay*ay);
    double t1, t2;
    if (h <= 2.0*ay) {
        double delta = h - ay;
        t1 = ax*(2.0*delta - ax);
        t2 = (delta - 2.0*(ax - ay))*delta;
    } else {
        double delta = h - ax;
        t1 = 2.0*delta*(ax - 2.0*ay);
        t2 = (4.0*delta - ay)*ay + delta*delta;
    }
    h -= (t1 + t2)/(2.0*h);
    return h;
}

BT_DEV double ref_hypot(double x, double y)
{
    if (!__builtin_isfinite(x) || !__builtin_isfinite(y)) {
        if (__builtin_isinf(x) || __builtin_isinf(y)) return __builtin_inf();
        return x + y;
    }
    x = fabs(x);
    y = fabs(y);
    const double ax = x < y ? y : x;
    const double ay = x < y ? x : y;
    constexpr double kScale = 0x1p-600, kEps = 0x1p-54;
    if (ax > 0x1p+511) {
        if (ay <= ax*kEps) return ax + ay;
        return hypot_kernel(ax*kScale, ay*kScale)/kScale;
    }
    if (ay < 0x1p-511) {
        if (ax >= ay/kEps) return ax + ay;
        return hypot_kernel(ax/kScale, ay/kScale)*kScale;
    }
    if (ay <= ax*kEps) return ax + ay;
    return hypot_kernel(ax, ay);
}

// The comparison hypot(x, y) < r decided from s = x*x + y*y without the
// square root: -1 when hypot(x, y) < r for certain, +1 when > r, 0 when s
// lies within 2^-43 (~1.1e-13) of r*r (or outside the range where that
// bound holds: NaN, inf, r outside {0} and (2^-465, 2^465), s >= 2^930),
// where only libm's value decides.  s carries at most ~2^-51 relative error
// and libm's hypot is within 1 ulp of the true value, far inside the 2^-43
// margin, so a decided comparison agrees with libm's.  The bounds are powers
// of two so that each compare takes its constant as an instruction literal
// instead of holding it in registers across the interface loop.
BT_DEV int cmp_hypot(double x, double y, double r)
{
    const double s = x*x + y*y;
    if (r == 0.0) return (s > 0.0 && s < 0x1p930) ? 1 : 0;   // an annulus from 0
    if (!(r > 0x1p-465 && r < 0x1p465)) return 0;
    if (!(s < 0x1p930)) return 0;
    const double r2 = r*r;
    const double d = s - r2;
    const double m = r2*0x1p-43;
    if (d > m) return 1;
    if (d < -m) return -1;
    return 0;
}

// ---------------------------------------------------------------------------
// Obscuration programs (src/obscuration.cpp); the stack is a bit mask.
// ---------------------------------------------------------------------------
BT_DEV bool polygon_contains(const bt_obsc* o, double x, double y)
{                                                    // obscuration.cpp:201-225
    const double* __restrict__ xp = o->a0;
    const double* __restrict__ yp = o->a1;
    const int size = o->n;
    double x1 = xp[0];
    double y1 = yp[0];
    double xinters = 0.0;
    bool inside = false;
    for (int i = 1; i <= size; i++) {
        const int k = (i == size) ? 0 : i;
        double x2 = xp[k];
        double y2 = yp[k];
        double ymin = (y2 < y1) ? y2 : y1;
        double ymax = (y1 < y2) ? y2 : y1;
        double xmax = (x1 < x2) ? x2 : x1;
        if (y > ymin && y <= ymax && x <= xmax) {
            if (y1 != y2)
                xinters = (y-y1)*(x2-x1)/(y2-y1)+x1;
            if (x1 == x2 || x <= xinters)
                inside = !inside;
        }
        x1 = x2;
        y1 = y2;
    }
    return inside;
}

BT_DEV bool obsc_contains(const bt_obsc* prog, int len, double x, double y)
{
    uint32_t stack = 0;   // bit k = stack slot k
    int sp = 0;
    for (int k = 0; k < len; k++) {
        const bt_obsc* o = prog + k;
        const double* p = o->p;
        bool v;
        switch (o->op) {
        case BT_OB_CIRCLE: {                         // obscuration.cpp:32-34
            const double dx = x-p[1], dy = y-p[2];
            const int c = cmp_hypot(dx, dy, p[0]);
            v = c ? (c < 0) : (ref_hypot(dx, dy) < p[0]);
            break;
        }
        case BT_OB_ANNULUS: {                        // obscuration.cpp:80-83
            const double dx = x-p[2], dy = y-p[3];
            const int ci = cmp_hypot(dx, dy, p[0]), co = cmp_hypot(dx, dy, p[1]);
            if (ci && co) {
                v = (ci > 0) && (co < 0);
            } else {
                const double h = ref_hypot(dx, dy);
                v = (p[0] <= h) && (h < p[1]);
            }
            break;
        }
        case BT_OB_RECTANGLE: {                      // obscuration.cpp:117-121
            double xp = (x-p[2])*p[4] + (y-p[3])*p[5];
            double yp = -(x-p[2])*p[5] + (y-p[3])*p[4];
            v = (xp > -p[0]/2 && xp < p[0]/2 && yp > -p[1]/2 && yp < p[1]/2);
            break;
        }
        case BT_OB_RAY: {                            // obscuration.cpp:155-159
            double xp = (x-p[1])*p[3] + (y-p[2])*p[4];
            double yp = -(x-p[1])*p[4] + (y-p[2])*p[3];
            v = (xp > 0.0 && yp > -p[0]/2 && yp < p[0]/2);
            break;
        }
        case BT_OB_POLYGON:
            v = polygon_contains(o, x, y);
            break;
        case BT_OB_NEGATION:                         // obscuration.cpp:320-322
            stack ^= (1u << (sp-1));
            continue;
        case BT_OB_UNION:                            // obscuration.cpp:364-370
        case BT_OB_INTERSECTION: {                   // obscuration.cpp:415-421
            const int m = o->n;
            const uint32_t mask = ((m >= 32) ? 0xffffffffu : ((1u << m) - 1u)) << (sp - m);
            const uint32_t bits = stack & mask;
            v = (o->op == BT_OB_UNION) ? (bits != 0) : (bits == mask);
            stack &= ~mask;
            sp -= m;
            break;
        }
        default:
            v = false;
        }
        stack = (stack & ~(1u << sp)) | ((uint32_t)v << sp);
        sp++;
    }
    return sp > 0 && ((stack >> (sp-1)) & 1u);
}

// ---------------------------------------------------------------------------
// One Interface: transform, intersect, interact, obscure
// (src/batoid.cpp:221-505 and optic.py:419-422)
// ---------------------------------------------------------------------------
struct Ray {
    double x, y, z, vx, vy, vz, t, w;
    bool vig, fail;
};

// Returns true when the ray's kinematic state was (re)written.  `n2` is the
// outgoing medium's index at the ray's wavelength (used by REFRACT only); the
// reference evaluates it after the intersection (batoid.cpp:474), but it is
// a pure function of the wavelength, so callers may evaluate it once per ray
// and medium and reuse it across Interfaces.
template <int K = kKindsAll>
BT_DEV bool apply_interaction(int action, const double* __restrict__ dr,
                              const double* __restrict__ drot,
                              const bt_surface* surf, double n2,
                              Ray& r, int niter, int start = BT_START_REFERENCE,
                              double correction = 0.0)
{
    if (r.fail) return false;
    double dx = r.x-dr[0];
    double dy = r.y-dr[1];
    double dz = r.z-dr[2];
    double x = dx*drot[0] + dy*drot[3] + dz*drot[6];
    double y = dx*drot[1] + dy*drot[4] + dz*drot[7];
    double z = dx*drot[2] + dy*drot[5] + dz*drot[8];
    double vx = r.vx*drot[0] + r.vy*drot[3] + r.vz*drot[6];
    double vy = r.vx*drot[1] + r.vy*drot[4] + r.vz*drot[7];
    double vz = r.vx*drot[2] + r.vy*drot[5] + r.vz*drot[8];
    double t = r.t;
    double dt = 0.0;
    NormalMemo memo;
    memo.ok = false;
    const bool hit = (start == BT_START_REFERENCE)
                         ? surf_tti<K>(surf, x, y, z, vx, vy, vz, dt, niter, &memo)
                         : fast_tti<K>(surf, x, y, z, vx, vy, vz, dt, niter, correction, &memo);
    // the normal at the intersection, reusing Newton's when it was evaluated
    // at this very point
    auto normal_at = [&](double px, double py, double& nx, double& ny, double& nz) {
        if (memo.ok && same_bits(memo.dt, dt)) {
            nx = memo.nx; ny = memo.ny; nz = memo.nz;
        } else {
            surf_normal<K>(surf, px, py, nx, ny, nz);
        }
    };
    if (!hit) {
        r.fail = true;
        r.vig = true;
        return false;          // left in the previous frame
    }
    x += vx * dt;
    y += vy * dt;
    z += vz * dt;
    t += dt;
    if (action == BT_ACT_REFLECT) {                  // batoid.cpp:340-353
        double nx, ny, nz;
        normal_at(x, y, nx, ny, nz);
        double alpha = vx*nx;
        alpha += vy*ny;
        alpha += vz*nz;
        vx -= 2*alpha*nx;
        vy -= 2*alpha*ny;
        vz -= 2*alpha*nz;
    } else if (action == BT_ACT_REFRACT) {           // batoid.cpp:453-496
        double n1 = vx*vx;
        n1 += vy*vy;
        n1 += vz*vz;
        n1 = cr_rsqrt(n1);
        double nvx = vx*n1;
        double nvy = vy*n1;
        double nvz = vz*n1;
        double nx, ny, nz;
        normal_at(x, y, nx, ny, nz);
        double alpha = nvx*nx;
        alpha += nvy*ny;
        alpha += nvz*nz;
        if (alpha > 0.) {
            nx *= -1; ny *= -1; nz *= -1;
            alpha *= -1;
        }
        // Into a medium of index exactly 1 (vacuum, or air given as a
        // constant 1) the reference's divisions by n2 return their numerators
        // bit for bit (x/1 == x in IEEE arithmetic, signed zeros, infinities
        // and NaNs included), so that case runs without them.
        auto finish = [&](auto unit) {
            constexpr bool kUnit = decltype(unit)::value;
            double eta = n1;
            if constexpr (!kUnit) eta /= n2;
            double sinsqr = eta*eta*(1-alpha*alpha);
            if (sinsqr > 1.0) {                      // total internal reflection
                vx = __builtin_nan(""); vy = vx; vz = vx;
                r.fail = true;
                r.vig = true;
            } else {
                double nfactor = eta*alpha + cr_sqrt(1-sinsqr);
                vx = eta*nvx - nfactor*nx;
                vy = eta*nvy - nfactor*ny;
                vz = eta*nvz - nfactor*nz;
                if constexpr (!kUnit) {
                    vx /= n2;
                    vy /= n2;
                    vz /= n2;
                }
            }
        };
        if (n2 == 1.0)
            finish(std::true_type{});
        else
            finish(std::false_type{});
    }
    r.x = x; r.y = y; r.z = z;
    r.vx = vx; r.vy = vy; r.vz = vz;
    r.t = t;
    return true;
}

}  // namespace bt
