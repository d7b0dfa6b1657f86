/*
 * oracle.c -- CPU restatement of jmeyers314/batoid's trace hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function cites the
 * reference file:line whose arithmetic it restates.  Operation order follows
 * the reference expression by expression, and the file is compiled with
 * -O2 -ffp-contract=off and no -march, as the reference's RelWithDebInfo
 * build for baseline x86-64 (setup.py:39, CMakeLists.txt:12), so results
 * are bit-identical to the reference's CPU trace.
 *
 * Parity pins (tests/test_oracle_*.py): bit-for-bit against the reference
 * C++ core compiled from its own sources into oracle/_ref; Zemax golden
 * traces (LSST at 1e-10 m); refractive-index CSVs at rtol 1e-13; closed
 * forms (paraboloid focus, vertical-ray sags, Snell, reflection law).
 */
#include "oracle.h"

#include <math.h>
#include <string.h>
#include <omp.h>

/* ------------------------------------------------------------------------
 * Surfaces
 * ---------------------------------------------------------------------- */

/* Table::eval, src/table.cpp:42-90 (oneDSpline at :42-49). */
static double spline1(double x, double v0, double v1, double d0, double d1)
{
    double a = 2*(v0-v1) + d0 + d1;
    double b = 3*(v1-v0) - 2*d0 - d1;
    double c = d0;
    double d = v0;
    return d + x*(c + x*(b + x*a));
}

/* oneDGrad, src/table.cpp:51-56. */
static double grad1(double x, double v0, double v1, double d0, double d1)
{
    double a = 2*(v0-v1) + d0 + d1;
    double b = 3*(v1-v0) - 2*d0 - d1;
    double c = d0;
    return c + x*(2*b + x*3*a);
}

/* Cell index as the reference computes it: int(std::floor(u)).  x86-64
 * converts NaN and out-of-range values to INT_MIN (cvttsd2si), which the
 * range test then rejects (src/table.cpp:59-63). */
static int cell_index(double u)
{
    double f = floor(u);
    if (!(f >= -2147483648.0 && f < 2147483648.0)) return (int)0x80000000;
    return (int)f;
}

static int table_cell(const bt_surface* s, double x, double y,
                      int* ix, int* iy)
{
    const double x0 = s->p[0], y0 = s->p[1], dx = s->p[2], dy = s->p[3];
    *ix = cell_index((x-x0)/dx);
    *iy = cell_index((y-y0)/dy);
    /* src/table.cpp:61: ix is compared against size_t nx-1, so negative
     * indices fail the first test as well; the combined effect is this. */
    return !(*ix >= s->n1-1 || *ix < 0 || *iy >= s->n2-1 || *iy < 0);
}

static double table_eval(const bt_surface* s, double x, double y)
{
    int ix, iy;
    if (!table_cell(s, x, y, &ix, &iy))
        return s->flags ? NAN : 0.0;
    const double x0 = s->p[0], y0 = s->p[1], dx = s->p[2], dy = s->p[3];
    const int64_t nx = s->n1;
    const double* z = s->a0; const double* zx = s->a1;
    const double* zy = s->a2; const double* zxy = s->a3;
    double xgrid = x0 + ix*dx;
    double ygrid = y0 + iy*dy;
    double xfrac = (x - xgrid)/dx;
    double yfrac = (y - ygrid)/dy;
    int64_t k0 = nx*iy + ix, k1 = nx*(iy+1) + ix;
    double val0 = spline1(xfrac, z[k0], z[k0+1], zx[k0]*dx, zx[k0+1]*dx);
    double val1 = spline1(xfrac, z[k1], z[k1+1], zx[k1]*dx, zx[k1+1]*dx);
    double der0 = spline1(xfrac, zy[k0], zy[k0+1], zxy[k0]*dx, zxy[k0+1]*dx);
    double der1 = spline1(xfrac, zy[k1], zy[k1+1], zxy[k1]*dx, zxy[k1+1]*dx);
    return spline1(yfrac, val0, val1, der0*dy, der1*dy);
}

/* Table::grad, src/table.cpp:92-158. */
static void table_grad(const bt_surface* s, double x, double y,
                       double* dzdx, double* dzdy)
{
    int ix, iy;
    if (!table_cell(s, x, y, &ix, &iy)) {
        double v = s->flags ? NAN : 0.0;
        *dzdx = v; *dzdy = v;
        return;
    }
    const double x0 = s->p[0], y0 = s->p[1], dx = s->p[2], dy = s->p[3];
    const int64_t nx = s->n1;
    const double* z = s->a0; const double* zx = s->a1;
    const double* zy = s->a2; const double* zxy = s->a3;
    double xgrid = x0 + ix*dx;
    double ygrid = y0 + iy*dy;
    double xfrac = (x - xgrid)/dx;
    double yfrac = (y - ygrid)/dy;
    int64_t k0 = nx*iy + ix, k1 = nx*(iy+1) + ix;

    double val0 = grad1(xfrac, z[k0], z[k0+1], zx[k0]*dx, zx[k0+1]*dx);
    double val1 = grad1(xfrac, z[k1], z[k1+1], zx[k1]*dx, zx[k1+1]*dx);
    double der0 = grad1(xfrac, zy[k0], zy[k0+1], zxy[k0]*dx, zxy[k0+1]*dx);
    double der1 = grad1(xfrac, zy[k1], zy[k1+1], zxy[k1]*dx, zxy[k1+1]*dx);
    *dzdx = spline1(yfrac, val0, val1, der0*dy, der1*dy)/dx;

    val0 = grad1(yfrac, z[k0], z[k1], zy[k0]*dy, zy[k1]*dy);
    val1 = grad1(yfrac, z[k0+1], z[k1+1], zy[k0+1]*dy, zy[k1+1]*dy);
    der0 = grad1(yfrac, zx[k0], zx[k1], zxy[k0]*dy, zxy[k1]*dy);
    der1 = grad1(yfrac, zx[k0+1], zx[k1+1], zxy[k0+1]*dy, zxy[k1+1]*dy);
    *dzdy = spline1(xfrac, val0, val1, der0*dx, der1*dx)/dy;
}

/* horner / horner2d, src/polynomialSurface.cpp:50-66. */
static double horner(double x, const double* c, int n)
{
    double r = 0.0;
    for (int i = n-1; i >= 0; i--) { r *= x; r += c[i]; }
    return r;
}

static double horner2d(double x, double y, const double* c, int nx, int ny)
{
    double r = 0.0;
    for (int i = ny-1; i >= 0; i--) { r *= x; r += horner(y, c + (int64_t)i*nx, nx); }
    return r;
}

/* Quadric::sag / _dzdr, src/quadric.cpp:18-25, 72-76. */
static double quadric_sag(const bt_surface* s, double x, double y)
{
    const double R = s->p[0], cp1RR = s->p[8];
    double r2 = x*x + y*y;
    if (R != 0)
        return r2/(R*(1.+sqrt(1.-r2*cp1RR)));
    return 0.0;
}

static double quadric_dzdr(const bt_surface* s, double r)
{
    const double R = s->p[0], cp1RR = s->p[8];
    if (R != 0.0)
        return r/(R*sqrt(1-r*r*cp1RR));
    return 0.0;
}

/* Asphere::_dzdr, src/asphere.cpp:64-73. */
static double asphere_dzdr(const bt_surface* s, double r)
{
    double result = quadric_dzdr(s, r);
    double rr = r*r;
    double rrr = r;
    for (int i = 0; i < s->n1; i++) {
        result += s->a1[i]*rrr;
        rrr *= rr;
    }
    return result;
}

/* Records of a surface: a Sum is followed by its children, each a leaf
 * record or a Sum with its own children; a Sum's n1 counts all records
 * below it, so its next sibling starts 1 + n1 records on. */
static int subtree_len(const bt_surface* s)
{
    return s->kind == BT_SUM ? 1 + s->n1 : 1;
}

double or_sag(const bt_surface* s, double x, double y)
{
    switch (s->kind) {
    case BT_PLANE:                                   /* plane.cpp:15-17 */
        return 0.0;
    case BT_SPHERE:                                  /* sphere.cpp:17-21 */
        if (s->p[0] != 0)
            return s->p[0]*(1-sqrt(1-(x*x + y*y)*s->p[3]));
        return 0.0;
    case BT_PARABOLOID: {                            /* paraboloid.cpp:17-23 */
        if (s->p[0] != 0) {
            double r2 = x*x + y*y;
            return r2*s->p[2];
        }
        return 0.0;
    }
    case BT_QUADRIC:                                 /* quadric.cpp:18-25 */
        return quadric_sag(s, x, y);
    case BT_ASPHERE: {                               /* asphere.cpp:37-46 */
        double r2 = x*x + y*y;
        double rr = 1.0;
        double result = quadric_sag(s, x, y);
        for (int i = 0; i < s->n1; i++) {
            rr *= r2;
            result += s->a0[i]*rr;
        }
        return result;
    }
    case BT_TILTED:                                  /* tilted.cpp:16-18 */
        return x*s->p[0] + y*s->p[1];
    case BT_BICUBIC:                                 /* bicubic.cpp:19-21 */
        return table_eval(s, x, y);
    case BT_POLYNOMIAL:                              /* polynomialSurface.cpp:35-37 */
        return horner2d(x-s->p[0], y-s->p[1], s->a0, s->n1, s->n2);
    case BT_SUM: {                                   /* sum.cpp:23-28 */
        double result = 0.0;
        const bt_surface* c = s + 1;
        for (int i = 0; i < s->nchild; i++, c += subtree_len(c))
            result += or_sag(c, x, y);
        return result;
    }
    }
    return NAN;
}

void or_normal(const bt_surface* s, double x, double y,
               double* nx, double* ny, double* nz)
{
    switch (s->kind) {
    case BT_PLANE:                                   /* plane.cpp:19-23 */
        *nx = 0.0; *ny = 0.0; *nz = 1.0;
        return;
    case BT_SPHERE: {                                /* sphere.cpp:23-35, 66-69 */
        const double R = s->p[0], Rinv = s->p[2];
        double rsqr = x*x + y*y;
        if (R == 0.0 || rsqr == 0.0) {
            *nx = 0.0; *ny = 0.0; *nz = 1.0;
        } else {
            double r = sqrt(rsqr);
            double rat = r*Rinv;
            double dzdr = rat/sqrt(1-rat*rat);
            double n_z = 1/sqrt(1+dzdr*dzdr);
            *nz = n_z;
            *nx = -x/r*dzdr*n_z;
            *ny = -y/r*dzdr*n_z;
        }
        return;
    }
    case BT_PARABOLOID: {                            /* paraboloid.cpp:25-37 */
        const double R = s->p[0], Rinv = s->p[1];
        if (R == 0) {
            *nx = 0.0; *ny = 0.0; *nz = 1.0;
        } else {
            double n_z = 1/sqrt(1+(x*x+y*y)*Rinv*Rinv);
            *nz = n_z;
            *nx = -x*Rinv*n_z;
            *ny = -y*Rinv*n_z;
        }
        return;
    }
    case BT_QUADRIC:                                 /* quadric.cpp:27-39 */
    case BT_ASPHERE: {                               /* asphere.cpp:48-62 */
        double r = sqrt(x*x + y*y);
        if (r == 0.0) {
            *nx = 0.0; *ny = 0.0; *nz = 1.0;
        } else {
            double dzdr = (s->kind == BT_QUADRIC) ? quadric_dzdr(s, r)
                                                  : asphere_dzdr(s, r);
            double n_z = 1/sqrt(1+dzdr*dzdr);
            *nz = n_z;
            *nx = -x/r*dzdr*n_z;
            *ny = -y/r*dzdr*n_z;
        }
        return;
    }
    case BT_TILTED: {                                /* tilted.cpp:20-27 */
        const double tx = s->p[0], ty = s->p[1];
        *nx = -tx;
        *ny = -ty;
        *nz = sqrt(1 - tx*tx - ty*ty);
        return;
    }
    case BT_BICUBIC: {                               /* bicubic.cpp:23-40 */
        double dxdz, dydz;
        table_grad(s, x, y, &dxdz, &dydz);
        if (isnan(dxdz)) {
            *nx = NAN; *ny = NAN; *nz = NAN;
            return;
        }
        double n_z = 1/sqrt(1 + dxdz*dxdz + dydz*dydz);
        *nz = n_z;
        *nx = -dxdz*n_z;
        *ny = -dydz*n_z;
        return;
    }
    case BT_POLYNOMIAL: {                            /* polynomialSurface.cpp:39-48 */
        double gx = -horner2d(x-s->p[0], y-s->p[1], s->a1, s->n1-1, s->n2-1);
        double gy = -horner2d(x-s->p[0], y-s->p[1], s->a2, s->n1-1, s->n2-1);
        double n_z = 1./sqrt(gx*gx + gy*gy + 1);
        *nz = n_z;
        *nx = gx*n_z;
        *ny = gy*n_z;
        return;
    }
    case BT_SUM: {                                   /* sum.cpp:30-45 */
        double sx = 0.0, sy = 0.0;
        const bt_surface* c = s + 1;
        for (int i = 0; i < s->nchild; i++, c += subtree_len(c)) {
            double tnx, tny, tnz;
            or_normal(c, x, y, &tnx, &tny, &tnz);
            sx += tnx/tnz;
            sy += tny/tnz;
        }
        double n_z = 1./sqrt(sx*sx + sy*sy + 1);
        *nz = n_z;
        *nx = sx*n_z;
        *ny = sy*n_z;
        return;
    }
    }
    *nx = NAN; *ny = NAN; *nz = NAN;
}

/* Surface::timeToIntersect, the tangent-plane Newton iteration with exactly
 * niter iterations and no early exit, src/surface.cpp:22-49. */
static int newton(const bt_surface* s, double x, double y, double z,
                  double vx, double vy, double vz, double* dtio, int niter)
{
    double dt = *dtio;
    double rPx = x+vx*dt;
    double rPy = y+vy*dt;
    double rPz = z+vz*dt;
    double sz = or_sag(s, rPx, rPy);
    for (int iter = 0; iter < niter; iter++) {
        double nx, ny, nz;
        or_normal(s, rPx, rPy, &nx, &ny, &nz);
        dt = (rPx-x)*nx + (rPy-y)*ny + (sz-z)*nz;
        dt /= (nx*vx + ny*vy + nz*vz);
        rPx = x+vx*dt;
        rPy = y+vy*dt;
        rPz = z+vz*dt;
        sz = or_sag(s, rPx, rPy);
    }
    *dtio = dt;
    return fabs(sz-rPz) < 1e-12;
}

/* Quadric::timeToIntersect, src/quadric.cpp:41-70. */
static int quadric_tti(const bt_surface* s, double x, double y, double z,
                       double vx, double vy, double vz, double* dt)
{
    const double cp1inv = s->p[5], Rcp1 = s->p[6], RRcp1cp1 = s->p[7];
    double z0term = z-Rcp1;
    double vrr0 = vx*x + vy*y;
    double a = vz*vz + (vx*vx+vy*vy)*cp1inv;
    double b = 2*(vz*z0term + vrr0*cp1inv);
    double c = z0term*z0term - RRcp1cp1 + (x*x + y*y)*cp1inv;
    double discriminant = b*b - 4*a*c;
    if (discriminant < 0)
        return 0;
    double dt1;
    if (b > 0)
        dt1 = (-b - sqrt(discriminant)) / (2*a);
    else
        dt1 = 2*c / (-b + sqrt(discriminant));
    double dt2 = c / (a*dt1);
    double z1 = z + vz*dt1;
    double z2 = z + vz*dt2;
    *dt = (fabs(z1) < fabs(z2)) ? dt1 : dt2;
    return 1;
}

int or_time_to_intersect(const bt_surface* s, double x, double y, double z,
                         double vx, double vy, double vz, double* dt,
                         int niter)
{
    switch (s->kind) {
    case BT_PLANE:                                   /* plane.cpp:25-31 */
        if (vz == 0)
            return 0;
        *dt = -z/vz;
        return 1;
    case BT_SPHERE: {                                /* sphere.cpp:37-64 */
        const double R = s->p[0], Rsq = s->p[1];
        double vr2 = vx*vx + vy*vy;
        double vz2 = vz*vz;
        double vrr0 = vx*x + vy*y;
        double r02 = x*x + y*y;
        double z0term = z-R;
        double a = vz2 + vr2;
        double b = 2*vz*z0term + 2*vrr0;
        double c = z0term*z0term - Rsq + r02;
        double discriminant = b*b - 4*a*c;
        if (discriminant < 0)
            return 0;
        double dt1;
        if (b > 0)
            dt1 = (-b - sqrt(discriminant)) / (2*a);
        else
            dt1 = 2*c / (-b + sqrt(discriminant));
        double dt2 = c / (a*dt1);
        /* std::max(a,b) = (a<b)?b:a ; std::min(a,b) = (b<a)?b:a */
        if (vz*R < 0)
            *dt = (dt1 < dt2) ? dt2 : dt1;
        else
            *dt = (dt2 < dt1) ? dt2 : dt1;
        return 1;
    }
    case BT_PARABOLOID: {                            /* paraboloid.cpp:39-71 */
        const double Rinv = s->p[1], tRinv = s->p[2], R = s->p[0];
        double a = (vx*vx + vy*vy)*tRinv;
        double b = (x*vx + y*vy)*Rinv - vz;
        double c = (x*x + y*y)*tRinv - z;
        if (a == 0) {
            *dt = -c/b;
            return 1;
        }
        double discriminant = b*b - 4*a*c;
        if (discriminant < 0)
            return 0;
        double dt1;
        if (b > 0)
            dt1 = (-b - sqrt(discriminant)) / (2*a);
        else
            dt1 = 2*c / (-b + sqrt(discriminant));
        double dt2 = c / (a*dt1);
        if (vz*R < 0)
            *dt = (dt1 < dt2) ? dt2 : dt1;
        else
            *dt = (dt2 < dt1) ? dt2 : dt1;
        return 1;
    }
    case BT_QUADRIC:
        return quadric_tti(s, x, y, z, vx, vy, vz, dt);
    case BT_ASPHERE:                                 /* asphere.cpp:75-85 */
        if (!quadric_tti(s, x, y, z, vx, vy, vz, dt))
            return 0;
        return newton(s, x, y, z, vx, vy, vz, dt, niter);
    case BT_TILTED: {                                /* tilted.cpp:29-37 */
        double num = z - x*s->p[0] - y*s->p[1];
        double den = vx*s->p[0] + vy*s->p[1] - vz;
        if (den == 0) return 0;
        *dt = num/den;
        return 1;
    }
    case BT_BICUBIC:                                 /* inherits surface.cpp:22 */
    case BT_POLYNOMIAL:
        return newton(s, x, y, z, vx, vy, vz, dt, niter);
    case BT_SUM:                                     /* sum.cpp:47-56 */
        if (!or_time_to_intersect(s + 1, x, y, z, vx, vy, vz, dt, niter))
            return 0;
        return newton(s, x, y, z, vx, vy, vz, dt, niter);
    }
    return 0;
}

/* ------------------------------------------------------------------------
 * Media, src/medium.cpp
 * ---------------------------------------------------------------------- */
double or_get_n(const bt_medium* m, double wavelength)
{
    switch (m->kind) {
    case BT_MED_CONST:                               /* medium.cpp:30-32 */
        return m->p[0];
    case BT_MED_TABLE: {                             /* medium.cpp:89-105 */
        const double* args = m->a0; const double* vals = m->a1;
        int size = m->n;
        if (wavelength < args[0]) return NAN;
        if (wavelength > args[size-1]) return NAN;
        int upperIdx;
        for (upperIdx = 1; upperIdx < size; upperIdx++)
            if (wavelength < args[upperIdx]) break;
        double out = (wavelength - args[upperIdx-1]);
        out *= (vals[upperIdx] - vals[upperIdx-1]);
        out /= (args[upperIdx] - args[upperIdx-1]);
        out += vals[upperIdx-1];
        return out;
    }
    case BT_MED_SELLMEIER: {                         /* medium.cpp:145-149 */
        const double* p = m->p;
        double x = wavelength*wavelength*1e12;
        return sqrt(1.0 + p[0]*x/(x-p[3]) + p[1]*x/(x-p[4]) + p[2]*x/(x-p[5]));
    }
    case BT_MED_SUMITA: {                            /* medium.cpp:186-191 */
        const double* A = m->p;
        double x = wavelength*wavelength*1e12;
        double y = 1./x;
        return sqrt(A[0] + A[1]*x + y*(A[2] + y*(A[3] + y*(A[4] + y*A[5]))));
    }
    case BT_MED_AIR: {                               /* medium.cpp:237-245 */
        const double P = m->p[0], T = m->p[1], W = m->p[2];
        double sigma_squared = 1e-12 / (wavelength*wavelength);
        double n_minus_one = (64.328 + (29498.1 / (146.0 - sigma_squared))
                              + (255.4 / (41.0 - sigma_squared))) * 1.e-6;
        n_minus_one *= P * (1.0 + (1.049 - 0.0157 * T) * 1.e-6 * P)
                       / (720.883 * (1.0 + 0.003661 * T));
        n_minus_one -= (0.0624 - 0.000680 * sigma_squared)/(1.0 + 0.003661 * T)
                       * W * 1.e-6;
        return 1+n_minus_one;
    }
    }
    return NAN;
}

/* ------------------------------------------------------------------------
 * Obscurations, src/obscuration.cpp
 * ---------------------------------------------------------------------- */
static int polygon_contains(const bt_obsc* o, double x, double y)
{                                                    /* obscuration.cpp:201-225 */
    const double* xp = o->a0; const double* yp = o->a1;
    int size = o->n;
    double x1 = xp[0];
    double y1 = yp[0];
    double xinters = 0.0;
    int inside = 0;
    for (int i = 1; i <= size; i++) {
        double x2 = xp[i % size];
        double y2 = yp[i % size];
        double ymin = (y2 < y1) ? y2 : y1;
        double ymax = (y1 < y2) ? y2 : y1;
        double xmax = (x1 < x2) ? x2 : x1;
        if (y > ymin) {
            if (y <= ymax) {
                if (x <= xmax) {
                    if (y1 != y2)
                        xinters = (y-y1)*(x2-x1)/(y2-y1)+x1;
                    if (x1 == x2 || x <= xinters)
                        inside = !inside;
                }
            }
        }
        x1 = x2;
        y1 = y2;
    }
    return inside;
}

int or_contains(const bt_obsc* prog, int len, double x, double y)
{
    int stack[BT_OBSC_MAX_STACK];
    int sp = 0;
    for (int k = 0; k < len; k++) {
        const bt_obsc* o = prog + k;
        const double* p = o->p;
        switch (o->op) {
        case BT_OB_CIRCLE:                           /* obscuration.cpp:32-34 */
            stack[sp++] = hypot(x-p[1], y-p[2]) < p[0];
            break;
        case BT_OB_ANNULUS: {                        /* obscuration.cpp:80-83 */
            double h = hypot(x-p[2], y-p[3]);
            stack[sp++] = (p[0] <= h) && (h < p[1]);
            break;
        }
        case BT_OB_RECTANGLE: {                      /* obscuration.cpp:117-121 */
            double xp = (x-p[2])*p[4] + (y-p[3])*p[5];
            double yp = -(x-p[2])*p[5] + (y-p[3])*p[4];
            stack[sp++] = (xp > -p[0]/2 && xp < p[0]/2 &&
                           yp > -p[1]/2 && yp < p[1]/2);
            break;
        }
        case BT_OB_RAY: {                            /* obscuration.cpp:155-159 */
            double xp = (x-p[1])*p[3] + (y-p[2])*p[4];
            double yp = -(x-p[1])*p[4] + (y-p[2])*p[3];
            stack[sp++] = (xp > 0.0 && yp > -p[0]/2 && yp < p[0]/2);
            break;
        }
        case BT_OB_POLYGON:
            stack[sp++] = polygon_contains(o, x, y);
            break;
        case BT_OB_NEGATION:                         /* obscuration.cpp:320-322 */
            stack[sp-1] = !stack[sp-1];
            break;
        case BT_OB_UNION: {                          /* obscuration.cpp:364-370 */
            int r = 0;
            for (int i = 0; i < o->n; i++) r |= stack[sp-o->n+i];
            sp -= o->n;
            stack[sp++] = r;
            break;
        }
        case BT_OB_INTERSECTION: {                   /* obscuration.cpp:415-421 */
            int r = 1;
            for (int i = 0; i < o->n; i++) r &= stack[sp-o->n+i];
            sp -= o->n;
            stack[sp++] = r;
            break;
        }
        }
    }
    return sp > 0 ? stack[sp-1] : 0;
}

/* ------------------------------------------------------------------------
 * Vectorised queries
 * ---------------------------------------------------------------------- */
void or_sag_n(const bt_surface* s, const double* x, const double* y,
              double* out, int64_t n, int nthreads)
{
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < n; i++) out[i] = or_sag(s, x[i], y[i]);
}

void or_normal_n(const bt_surface* s, const double* x, const double* y,
                 double* out, int64_t n, int nthreads)
{                                                    /* pysrc/surface.cpp:12-21 */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < n; i++)
        or_normal(s, x[i], y[i], &out[i], &out[i+n], &out[i+2*n]);
}

void or_get_n_n(const bt_medium* m, const double* w, double* out, int64_t n,
                int nthreads)
{
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < n; i++) out[i] = or_get_n(m, w[i]);
}

void or_contains_n(const bt_obsc* prog, int len, const double* x,
                   const double* y, uint8_t* out, int64_t n, int nthreads)
{
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < n; i++) out[i] = (uint8_t)or_contains(prog, len, x[i], y[i]);
}

/* ------------------------------------------------------------------------
 * Transforms, src/batoid.cpp:8-148
 * ---------------------------------------------------------------------- */
void or_apply_forward_transform_arrays(const double* dr, const double* drot,
                                       double* x, double* y, double* z,
                                       int64_t n, int nthreads)
{                                                    /* batoid.cpp:8-22 */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < n; i++) {
        double dx = x[i]-dr[0];
        double dy = y[i]-dr[1];
        double dz = z[i]-dr[2];
        x[i] = dx*drot[0] + dy*drot[3] + dz*drot[6];
        y[i] = dx*drot[1] + dy*drot[4] + dz*drot[7];
        z[i] = dx*drot[2] + dy*drot[5] + dz*drot[8];
    }
}

void or_apply_reverse_transform_arrays(const double* dr, const double* drot,
                                       double* x, double* y, double* z,
                                       int64_t n, int nthreads)
{                                                    /* batoid.cpp:24-38 */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < n; i++) {
        double xx = x[i]*drot[0] + y[i]*drot[1] + z[i]*drot[2] + dr[0];
        double yy = x[i]*drot[3] + y[i]*drot[4] + z[i]*drot[5] + dr[1];
        double zz = x[i]*drot[6] + y[i]*drot[7] + z[i]*drot[8] + dr[2];
        x[i] = xx; y[i] = yy; z[i] = zz;
    }
}

void or_finish_parallel(const double* dr, const double* drot, const double* vv,
                        double* x, double* y, double* z, int64_t n)
{                                                    /* batoid.cpp:40-69 */
    double vxlocal = -vv[0]*drot[0] - vv[1]*drot[3] - vv[2]*drot[6];
    double vylocal = -vv[0]*drot[1] - vv[1]*drot[4] - vv[2]*drot[7];
    double vzlocal = -vv[0]*drot[2] - vv[1]*drot[5] - vv[2]*drot[8];
    for (int64_t i = 0; i < n; i++) {
        double dx = x[i]-dr[0];
        double dy = y[i]-dr[1];
        double dz = z[i]-dr[2];
        x[i] = dx*drot[0] + dy*drot[3] + dz*drot[6];
        y[i] = dx*drot[1] + dy*drot[4] + dz*drot[7];
        z[i] = dx*drot[2] + dy*drot[5] + dz*drot[8];
        double dt = -z[i]/vzlocal;
        x[i] += dt*vxlocal;
        y[i] += dt*vylocal;
        z[i] += dt*vzlocal;
        double xx = x[i]*drot[0] + y[i]*drot[1] + z[i]*drot[2] + dr[0];
        double yy = x[i]*drot[3] + y[i]*drot[4] + z[i]*drot[5] + dr[1];
        double zz = x[i]*drot[6] + y[i]*drot[7] + z[i]*drot[8] + dr[2];
        x[i] = xx; y[i] = yy; z[i] = zz;
    }
}

void or_apply_forward_transform(const double* dr, const double* drot,
                                bt_rays rv, int nthreads)
{                                                    /* batoid.cpp:71-108 */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < rv.n; i++) {
        double dx = rv.x[i]-dr[0];
        double dy = rv.y[i]-dr[1];
        double dz = rv.z[i]-dr[2];
        rv.x[i] = dx*drot[0] + dy*drot[3] + dz*drot[6];
        rv.y[i] = dx*drot[1] + dy*drot[4] + dz*drot[7];
        rv.z[i] = dx*drot[2] + dy*drot[5] + dz*drot[8];
        double vx = rv.vx[i]*drot[0] + rv.vy[i]*drot[3] + rv.vz[i]*drot[6];
        double vy = rv.vx[i]*drot[1] + rv.vy[i]*drot[4] + rv.vz[i]*drot[7];
        double vz = rv.vx[i]*drot[2] + rv.vy[i]*drot[5] + rv.vz[i]*drot[8];
        rv.vx[i] = vx; rv.vy[i] = vy; rv.vz[i] = vz;
    }
}

void or_apply_reverse_transform(const double* dr, const double* drot,
                                bt_rays rv, int nthreads)
{                                                    /* batoid.cpp:111-148 */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < rv.n; i++) {
        double x = rv.x[i]*drot[0] + rv.y[i]*drot[1] + rv.z[i]*drot[2] + dr[0];
        double y = rv.x[i]*drot[3] + rv.y[i]*drot[4] + rv.z[i]*drot[5] + dr[1];
        double z = rv.x[i]*drot[6] + rv.y[i]*drot[7] + rv.z[i]*drot[8] + dr[2];
        rv.x[i] = x; rv.y[i] = y; rv.z[i] = z;
        double vx = rv.vx[i]*drot[0] + rv.vy[i]*drot[1] + rv.vz[i]*drot[2];
        double vy = rv.vx[i]*drot[3] + rv.vy[i]*drot[4] + rv.vz[i]*drot[5];
        double vz = rv.vx[i]*drot[6] + rv.vy[i]*drot[7] + rv.vz[i]*drot[8];
        rv.vx[i] = vx; rv.vy[i] = vy; rv.vz[i] = vz;
    }
}

/* ------------------------------------------------------------------------
 * Surface interactions, src/batoid.cpp:175-506 (coating = nullptr)
 * ---------------------------------------------------------------------- */
#define LOCAL_FRAME                                                        \
    double dx = rv.x[i]-dr[0];                                             \
    double dy = rv.y[i]-dr[1];                                             \
    double dz = rv.z[i]-dr[2];                                             \
    double x = dx*drot[0] + dy*drot[3] + dz*drot[6];                       \
    double y = dx*drot[1] + dy*drot[4] + dz*drot[7];                       \
    double z = dx*drot[2] + dy*drot[5] + dz*drot[8];                       \
    double vx = rv.vx[i]*drot[0] + rv.vy[i]*drot[3] + rv.vz[i]*drot[6];    \
    double vy = rv.vx[i]*drot[1] + rv.vy[i]*drot[4] + rv.vz[i]*drot[7];    \
    double vz = rv.vx[i]*drot[2] + rv.vy[i]*drot[5] + rv.vz[i]*drot[8];    \
    double t = rv.t[i];

/* coat: NULL, or the coating's flux factor for rays that reach the
 * surface (batoid.cpp:251-263, 362-371, 497-499; SimpleCoating's factors
 * are constants, coating.cpp:28-39). */
static void intersect_c(const bt_surface* s, const double* dr, const double* drot,
                        bt_rays rv, int niter, int nthreads, const double* coat)
{                                                    /* batoid.cpp:221-269 */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < rv.n; i++) {
        LOCAL_FRAME
        if (!rv.failed[i]) {
            double dt = 0.0;
            int success = or_time_to_intersect(s, x, y, z, vx, vy, vz, &dt, niter);
            if (success) {
                x += vx * dt;
                y += vy * dt;
                z += vz * dt;
                t += dt;
                rv.x[i] = x; rv.y[i] = y; rv.z[i] = z;
                rv.vx[i] = vx; rv.vy[i] = vy; rv.vz[i] = vz;
                rv.t[i] = t;
                if (coat) rv.flux[i] *= *coat;
            } else {
                rv.failed[i] = 1;
                rv.vignetted[i] = 1;
            }
        }
    }
}

void or_intersect(const bt_surface* s, const double* dr, const double* drot,
                  bt_rays rv, int niter, int nthreads)
{
    intersect_c(s, dr, drot, rv, niter, nthreads, NULL);
}

void or_intersect_coated(const bt_surface* s, const double* dr, const double* drot,
                         double transmissivity, bt_rays rv, int niter, int nthreads)
{
    intersect_c(s, dr, drot, rv, niter, nthreads, &transmissivity);
}

static void reflect_c(const bt_surface* s, const double* dr, const double* drot,
                      bt_rays rv, int niter, int nthreads, const double* coat)
{                                                    /* batoid.cpp:319-377 */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < rv.n; i++) {
        LOCAL_FRAME
        if (!rv.failed[i]) {
            double dt = 0.0;
            int success = or_time_to_intersect(s, x, y, z, vx, vy, vz, &dt, niter);
            if (success) {
                x += vx * dt;
                y += vy * dt;
                z += vz * dt;
                t += dt;
                double nx, ny, nz;
                or_normal(s, x, y, &nx, &ny, &nz);
                double alpha = vx*nx;
                alpha += vy*ny;
                alpha += vz*nz;
                vx -= 2*alpha*nx;
                vy -= 2*alpha*ny;
                vz -= 2*alpha*nz;
                rv.x[i] = x; rv.y[i] = y; rv.z[i] = z;
                rv.vx[i] = vx; rv.vy[i] = vy; rv.vz[i] = vz;
                rv.t[i] = t;
                if (coat) rv.flux[i] *= *coat;
            } else {
                rv.failed[i] = 1;
                rv.vignetted[i] = 1;
            }
        }
    }
}

void or_reflect(const bt_surface* s, const double* dr, const double* drot,
                bt_rays rv, int niter, int nthreads)
{
    reflect_c(s, dr, drot, rv, niter, nthreads, NULL);
}

void or_reflect_coated(const bt_surface* s, const double* dr, const double* drot,
                       double reflectivity, bt_rays rv, int niter, int nthreads)
{
    reflect_c(s, dr, drot, rv, niter, nthreads, &reflectivity);
}

static void refract_c(const bt_surface* s, const double* dr, const double* drot,
                      const bt_medium* m2, bt_rays rv, int niter, int nthreads,
                      const double* coat)
{                                                    /* batoid.cpp:429-505 */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < rv.n; i++) {
        LOCAL_FRAME
        if (!rv.failed[i]) {
            double dt = 0.0;
            int success = or_time_to_intersect(s, x, y, z, vx, vy, vz, &dt, niter);
            if (success) {
                x += vx * dt;
                y += vy * dt;
                z += vz * dt;
                t += dt;
                double n1 = vx*vx;
                n1 += vy*vy;
                n1 += vz*vz;
                n1 = 1/sqrt(n1);
                double nvx = vx*n1;
                double nvy = vy*n1;
                double nvz = vz*n1;
                double nx, ny, nz;
                or_normal(s, x, y, &nx, &ny, &nz);
                double alpha = nvx*nx;
                alpha += nvy*ny;
                alpha += nvz*nz;
                if (alpha > 0.) {
                    nx *= -1; ny *= -1; nz *= -1;
                    alpha *= -1;
                }
                double n2 = or_get_n(m2, rv.wavelength[i]);
                double eta = n1/n2;
                double sinsqr = eta*eta*(1-alpha*alpha);
                rv.x[i] = x; rv.y[i] = y; rv.z[i] = z;
                rv.t[i] = t;
                if (sinsqr > 1.0) {
                    rv.vx[i] = NAN; rv.vy[i] = NAN; rv.vz[i] = NAN;
                    rv.failed[i] = 1;
                    rv.vignetted[i] = 1;
                } else {
                    double nfactor = eta*alpha + sqrt(1-sinsqr);
                    double ovx = eta*nvx - nfactor*nx;
                    double ovy = eta*nvy - nfactor*ny;
                    double ovz = eta*nvz - nfactor*nz;
                    rv.vx[i] = ovx / n2;
                    rv.vy[i] = ovy / n2;
                    rv.vz[i] = ovz / n2;
                }
                if (coat) rv.flux[i] *= *coat;
            } else {
                rv.failed[i] = 1;
                rv.vignetted[i] = 1;
            }
        }
    }
}

void or_refract(const bt_surface* s, const double* dr, const double* drot,
                const bt_medium* m2, bt_rays rv, int niter, int nthreads)
{
    refract_c(s, dr, drot, m2, rv, niter, nthreads, NULL);
}

void or_refract_coated(const bt_surface* s, const double* dr, const double* drot,
                       const bt_medium* m2, double transmissivity, bt_rays rv, int niter,
                       int nthreads)
{
    refract_c(s, dr, drot, m2, rv, niter, nthreads, &transmissivity);
}

void or_rsplit(const bt_surface* s, const double* dr, const double* drot,
               const bt_medium* m2, double reflectivity, double transmissivity,
               bt_rays rv, bt_rays rvs, int niter, int nthreads)
{                                                    /* batoid.cpp:509-674 */
    /* rvs holds a copy of rv on entry (trace.py:125); rays already failed
     * are left as they are in both */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < rv.n; i++) {
        LOCAL_FRAME
        if (!rv.failed[i]) {
            double dt = 0.0;
            int success = or_time_to_intersect(s, x, y, z, vx, vy, vz, &dt, niter);
            if (success) {
                x += vx * dt;
                y += vy * dt;
                z += vz * dt;
                t += dt;
                double n1 = vx*vx;
                n1 += vy*vy;
                n1 += vz*vz;
                n1 = 1/sqrt(n1);
                double nvx = vx*n1;
                double nvy = vy*n1;
                double nvz = vz*n1;
                double nx, ny, nz;
                or_normal(s, x, y, &nx, &ny, &nz);
                double alpha = nvx*nx;
                alpha += nvy*ny;
                alpha += nvz*nz;
                if (alpha > 0) {
                    nx *= -1; ny *= -1; nz *= -1;
                    alpha *= -1;
                }
                /* reflected part */
                rvs.x[i] = x; rvs.y[i] = y; rvs.z[i] = z;
                rvs.vx[i] = vx - 2*alpha*nx/n1;
                rvs.vy[i] = vy - 2*alpha*ny/n1;
                rvs.vz[i] = vz - 2*alpha*nz/n1;
                rvs.t[i] = t;
                rvs.wavelength[i] = rv.wavelength[i];
                rvs.flux[i] = rv.flux[i]*reflectivity;
                rvs.vignetted[i] = rv.vignetted[i];
                rvs.failed[i] = rv.failed[i];
                /* refracted part */
                double n2 = or_get_n(m2, rv.wavelength[i]);
                double eta = n1/n2;
                double sinsqr = eta*eta*(1-alpha*alpha);
                rv.x[i] = x; rv.y[i] = y; rv.z[i] = z;
                rv.t[i] = t;
                if (sinsqr > 1.0) {
                    rv.vx[i] = NAN; rv.vy[i] = NAN; rv.vz[i] = NAN;
                    rv.failed[i] = 1;
                    rv.vignetted[i] = 1;
                } else {
                    double nfactor = eta*alpha + sqrt(1-sinsqr);
                    double ovx = eta*nvx - nfactor*nx;
                    double ovy = eta*nvy - nfactor*ny;
                    double ovz = eta*nvz - nfactor*nz;
                    rv.vx[i] = ovx / n2;
                    rv.vy[i] = ovy / n2;
                    rv.vz[i] = ovz / n2;
                }
                rv.flux[i] *= transmissivity;
            } else {
                rv.vignetted[i] = 1;
                rv.failed[i] = 1;
                rvs.vignetted[i] = 1;
                rvs.failed[i] = 1;
            }
        }
    }
}

void or_obscure(const bt_obsc* prog, int len, bt_rays rv, int nthreads)
{                                                    /* batoid.cpp:151-172 */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < rv.n; i++)
        rv.vignetted[i] |= (uint8_t)or_contains(prog, len, rv.x[i], rv.y[i]);
}

void or_trace_plan(const bt_plan* plan, bt_rays rays, int niter, int nthreads)
{                                                    /* optic.py:391-424, 1013-1049 */
    for (int k = 0; k < plan->nstep; k++) {
        const bt_step* st = plan->steps + k;
        const bt_surface* s = plan->surfaces + st->surface;
        switch (st->action) {
        case BT_ACT_INTERSECT:
            or_intersect(s, st->dr, st->drot, rays, niter, nthreads); break;
        case BT_ACT_REFLECT:
            or_reflect(s, st->dr, st->drot, rays, niter, nthreads); break;
        case BT_ACT_REFRACT:
            or_refract(s, st->dr, st->drot, plan->media + st->medium, rays,
                       niter, nthreads);
            break;
        }
        if (st->obsc >= 0 && st->obsc_len > 0)
            or_obscure(plan->obscs + st->obsc, st->obsc_len, rays, nthreads);
    }
}

/* ------------------------------------------------------------------------
 * RayVector field operations, src/rayVector.cpp:26-222
 * ---------------------------------------------------------------------- */
static const double BT_PI = 3.14159265358979323846;

void or_position_at_time(bt_rays rv, double t, double* xo, double* yo,
                         double* zo, int nthreads)
{                                                    /* rayVector.cpp:26-54 */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < rv.n; i++) {
        xo[i] = rv.x[i] + rv.vx[i] * (t-rv.t[i]);
        yo[i] = rv.y[i] + rv.vy[i] * (t-rv.t[i]);
        zo[i] = rv.z[i] + rv.vz[i] * (t-rv.t[i]);
    }
}

void or_propagate_in_place(bt_rays rv, double t, int nthreads)
{                                                    /* rayVector.cpp:56-82 */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < rv.n; i++) {
        rv.x[i] += rv.vx[i] * (t - rv.t[i]);
        rv.y[i] += rv.vy[i] * (t - rv.t[i]);
        rv.z[i] += rv.vz[i] * (t - rv.t[i]);
        rv.t[i] = t;
    }
}

static double ray_phase(bt_rays rv, int64_t i, double x, double y, double z,
                        double t)
{                                                    /* rayVector.cpp:113-120 */
    double v2 = rv.vx[i]*rv.vx[i] + rv.vy[i]*rv.vy[i] + rv.vz[i]*rv.vz[i];
    double out = (x-rv.x[i])*rv.vx[i];
    out += (y-rv.y[i])*rv.vy[i];
    out += (z-rv.z[i])*rv.vz[i];
    out /= v2;
    out -= t-rv.t[i];
    out *= 2 * BT_PI / rv.wavelength[i];
    return out;
}

void or_phase(bt_rays rv, double x, double y, double z, double t,
              double* out, int nthreads)
{
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < rv.n; i++) out[i] = ray_phase(rv, i, x, y, z, t);
}

void or_amplitude(bt_rays rv, double x, double y, double z, double t,
                  double* out, int nthreads)
{                                                    /* rayVector.cpp:124-166 */
    #pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < rv.n; i++) {
        double ph = ray_phase(rv, i, x, y, z, t);
        out[2*i] = cos(ph);
        out[2*i+1] = sin(ph);
    }
}

void or_sum_amplitude(bt_rays rv, double x, double y, double z, double t,
                      int ignore_vignetted, double* out, int nthreads)
{                                                    /* rayVector.cpp:169-222 */
    double re = 0, im = 0;
    #pragma omp parallel for reduction(+:re, im) num_threads(nthreads)
    for (int64_t i = 0; i < rv.n; i++) {
        double ph = ray_phase(rv, i, x, y, z, t);
        if (!rv.failed[i] && !(ignore_vignetted && rv.vignetted[i])) {
            re += cos(ph)*rv.flux[i];
            im += sin(ph)*rv.flux[i];
        }
    }
    out[0] = re; out[1] = im;
}

void or_huygens_direct(bt_rays rv, const double* p0, const double* a1,
                       const double* a2, int nx, int ny, double t,
                       int ignore_vignetted, double* amp_out, int nthreads)
{                                                    /* analysis.py:296-303 */
    #pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int64_t pix = 0; pix < (int64_t)nx*ny; pix++) {
        int j = (int)(pix / nx), i = (int)(pix % nx);
        /* lattice.py:6-14 coords, transposed at analysis.py:296-297 */
        double ci = (double)(i - nx/2), cj = (double)(j - ny/2);
        double px = ci*a1[0] + cj*a2[0] + p0[0];
        double py = ci*a1[1] + cj*a2[1] + p0[1];
        double re = 0, im = 0;
        for (int64_t r = 0; r < rv.n; r++) {
            double ph = ray_phase(rv, r, px, py, 0.0, t);
            if (!rv.failed[r] && !(ignore_vignetted && rv.vignetted[r])) {
                re += cos(ph)*rv.flux[r];
                im += sin(ph)*rv.flux[r];
            }
        }
        amp_out[2*pix] = re; amp_out[2*pix+1] = im;
    }
}
