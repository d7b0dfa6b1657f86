// ref_harness.cpp -- drives the reference's own C++ trace core.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference sources
// (jmeyers314/batoid src/*.cpp, read in place under /root/reference by
// oracle/Makefile) into oracle/_ref/libbatoid_ref.so.  It builds reference
// objects (batoid::Sphere, batoid::Asphere, ...) from this project's POD
// descriptors and calls batoid::intersect / reflect / refract / obscure /
// RayVector::sumAmplitude, so the C restatement in oracle.c (and, through
// it, the HIP kernels) can be checked bit-for-bit against the reference.
#include <memory>
#include <vector>
#include <complex>

#include "batoid.h"
#include "surface.h"
#include "plane.h"
#include "sphere.h"
#include "paraboloid.h"
#include "quadric.h"
#include "asphere.h"
#include "tilted.h"
#include "bicubic.h"
#include "table.h"
#include "polynomialSurface.h"
#include "sum.h"
#include "medium.h"
#include "obscuration.h"
#include "rayVector.h"
#include "coating.h"

#include "../include/batoid_amd.h"

namespace {

struct Built {
    std::vector<std::unique_ptr<batoid::Surface>> surfaces;
    std::vector<std::unique_ptr<batoid::Table>> tables;
    std::vector<std::unique_ptr<const batoid::Surface*[]>> lists;
};

const batoid::Surface* build_leaf(const bt_surface* s, Built& b)
{
    batoid::Surface* out = nullptr;
    switch (s->kind) {
    case BT_PLANE: out = new batoid::Plane(); break;
    case BT_SPHERE: out = new batoid::Sphere(s->p[0]); break;
    case BT_PARABOLOID: out = new batoid::Paraboloid(s->p[0]); break;
    case BT_QUADRIC: out = new batoid::Quadric(s->p[0], s->p[1]); break;
    case BT_ASPHERE:
        out = new batoid::Asphere(s->p[0], s->p[1], s->a0, (size_t)s->n1);
        break;
    case BT_TILTED: out = new batoid::Tilted(s->p[0], s->p[1]); break;
    case BT_BICUBIC: {
        auto* t = new batoid::Table(
            s->p[0], s->p[1], s->p[2], s->p[3], s->a0, s->a1, s->a2, s->a3,
            (size_t)s->n1, (size_t)s->n2, s->flags != 0);
        b.tables.emplace_back(t);
        out = new batoid::Bicubic(t);
        break;
    }
    case BT_POLYNOMIAL:
        out = new batoid::PolynomialSurface(
            s->a0, s->a1, s->a2, s->p[0], s->p[1], (size_t)s->n1, (size_t)s->n2);
        break;
    default: return nullptr;
    }
    b.surfaces.emplace_back(out);
    return out;
}

// A Sum's children follow it, each a leaf record or a Sum with its n1
// descendant records after it.
const batoid::Surface* build_surface(const bt_surface* s, Built& b)
{
    if (s->kind != BT_SUM) return build_leaf(s, b);
    std::unique_ptr<const batoid::Surface*[]> kids(new const batoid::Surface*[s->nchild]);
    const bt_surface* c = s + 1;
    for (int i = 0; i < s->nchild; i++) {
        kids[i] = build_surface(c, b);
        c += (c->kind == BT_SUM) ? 1 + c->n1 : 1;
    }
    auto* sum = new batoid::Sum(kids.get(), (size_t)s->nchild);
    b.lists.push_back(std::move(kids));
    b.surfaces.emplace_back(sum);
    return sum;
}

std::unique_ptr<batoid::Medium> build_medium(const bt_medium* m)
{
    switch (m->kind) {
    case BT_MED_CONST: return std::unique_ptr<batoid::Medium>(new batoid::ConstMedium(m->p[0]));
    case BT_MED_TABLE:
        return std::unique_ptr<batoid::Medium>(
            new batoid::TableMedium(m->a0, m->a1, (size_t)m->n));
    case BT_MED_SELLMEIER:
        return std::unique_ptr<batoid::Medium>(new batoid::SellmeierMedium(
            m->p[0], m->p[1], m->p[2], m->p[3], m->p[4], m->p[5]));
    case BT_MED_SUMITA:
        return std::unique_ptr<batoid::Medium>(new batoid::SumitaMedium(
            m->p[0], m->p[1], m->p[2], m->p[3], m->p[4], m->p[5]));
    case BT_MED_AIR:
        // p[3..5] keep the constructor's SI inputs.
        return std::unique_ptr<batoid::Medium>(new batoid::Air(m->p[3], m->p[4], m->p[5]));
    }
    return nullptr;
}

struct ObscBuilt {
    std::vector<std::unique_ptr<batoid::Obscuration>> nodes;
    std::vector<std::unique_ptr<const batoid::Obscuration*[]>> lists;
};

const batoid::Obscuration* build_obsc(const bt_obsc* prog, int len, ObscBuilt& ob)
{
    std::vector<const batoid::Obscuration*> stack;
    for (int k = 0; k < len; k++) {
        const bt_obsc* o = prog + k;
        batoid::Obscuration* node = nullptr;
        switch (o->op) {
        case BT_OB_CIRCLE: node = new batoid::ObscCircle(o->p[0], o->p[1], o->p[2]); break;
        case BT_OB_ANNULUS:
            node = new batoid::ObscAnnulus(o->p[0], o->p[1], o->p[2], o->p[3]); break;
        case BT_OB_RECTANGLE:
            node = new batoid::ObscRectangle(o->p[0], o->p[1], o->p[2], o->p[3], o->p[6]);
            break;
        case BT_OB_RAY:
            node = new batoid::ObscRay(o->p[0], o->p[5], o->p[1], o->p[2]); break;
        case BT_OB_POLYGON: node = new batoid::ObscPolygon(o->a0, o->a1, (size_t)o->n); break;
        case BT_OB_NEGATION: {
            const batoid::Obscuration* a = stack.back(); stack.pop_back();
            node = new batoid::ObscNegation(a);
            break;
        }
        case BT_OB_UNION:
        case BT_OB_INTERSECTION: {
            std::unique_ptr<const batoid::Obscuration*[]> kids(new const batoid::Obscuration*[o->n]);
            for (int i = 0; i < o->n; i++) kids[i] = stack[stack.size() - o->n + i];
            stack.resize(stack.size() - o->n);
            if (o->op == BT_OB_UNION) node = new batoid::ObscUnion(kids.get(), (size_t)o->n);
            else node = new batoid::ObscIntersection(kids.get(), (size_t)o->n);
            ob.lists.push_back(std::move(kids));
            break;
        }
        }
        ob.nodes.emplace_back(node);
        stack.push_back(node);
    }
    return stack.empty() ? nullptr : stack.back();
}

batoid::RayVector make_rv(const bt_rays& r)
{
    return batoid::RayVector(r.x, r.y, r.z, r.vx, r.vy, r.vz, r.t, r.wavelength,
                             r.flux, reinterpret_cast<bool*>(r.vignetted),
                             reinterpret_cast<bool*>(r.failed), (size_t)r.n);
}

batoid::vec3 v3(const double* a) { return {a[0], a[1], a[2]}; }
batoid::mat3 m3(const double* a)
{
    return {a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8]};
}

}  // namespace

extern "C" {

void ref_intersect(const bt_surface* s, const double* dr, const double* drot,
                   bt_rays rays, int niter, int nthreads)
{
    Built b;
    const batoid::Surface* surf = build_surface(s, b);
    batoid::RayVector rv = make_rv(rays);
    batoid::intersect(*surf, v3(dr), m3(drot), rv, nullptr, nthreads, niter);
}

void ref_reflect(const bt_surface* s, const double* dr, const double* drot,
                 bt_rays rays, int niter, int nthreads)
{
    Built b;
    const batoid::Surface* surf = build_surface(s, b);
    batoid::RayVector rv = make_rv(rays);
    batoid::reflect(*surf, v3(dr), m3(drot), rv, nullptr, nthreads, niter);
}

void ref_refract(const bt_surface* s, const double* dr, const double* drot,
                 const bt_medium* m2, bt_rays rays, int niter, int nthreads)
{
    Built b;
    const batoid::Surface* surf = build_surface(s, b);
    auto med = build_medium(m2);
    batoid::RayVector rv = make_rv(rays);
    // m1 is unused by the reference kernel (n1 comes from |v|).
    batoid::refract(*surf, v3(dr), m3(drot), *med, *med, rv, nullptr, nthreads, niter);
}

// action 0/1/2 = intersect / reflect / refract with a SimpleCoating.
void ref_coated(int action, const bt_surface* s, const double* dr, const double* drot,
                const bt_medium* m2, double reflectivity, double transmissivity,
                bt_rays rays, int niter, int nthreads)
{
    Built b;
    const batoid::Surface* surf = build_surface(s, b);
    batoid::SimpleCoating coating(reflectivity, transmissivity);
    batoid::RayVector rv = make_rv(rays);
    if (action == 0) {
        batoid::intersect(*surf, v3(dr), m3(drot), rv, &coating, nthreads, niter);
    } else if (action == 1) {
        batoid::reflect(*surf, v3(dr), m3(drot), rv, &coating, nthreads, niter);
    } else {
        auto med = build_medium(m2);
        batoid::refract(*surf, v3(dr), m3(drot), *med, *med, rv, &coating, nthreads, niter);
    }
}

// rSplit; `split` must hold a copy of `rays` on entry (trace.py:125).
void ref_rsplit(const bt_surface* s, const double* dr, const double* drot,
                const bt_medium* m2, double reflectivity, double transmissivity,
                bt_rays rays, bt_rays split, int niter, int nthreads)
{
    Built b;
    const batoid::Surface* surf = build_surface(s, b);
    auto med = build_medium(m2);
    batoid::SimpleCoating coating(reflectivity, transmissivity);
    batoid::RayVector rv = make_rv(rays);
    batoid::RayVector rvs = make_rv(split);
    batoid::rSplit(*surf, v3(dr), m3(drot), *med, *med, coating, rv, rvs, nthreads, niter);
}

void ref_obscure(const bt_obsc* prog, int len, bt_rays rays, int nthreads)
{
    ObscBuilt ob;
    const batoid::Obscuration* o = build_obsc(prog, len, ob);
    batoid::RayVector rv = make_rv(rays);
    batoid::obscure(*o, rv, nthreads);
}

void ref_sag_n(const bt_surface* s, const double* x, const double* y,
               double* out, int64_t n)
{
    Built b;
    const batoid::Surface* surf = build_surface(s, b);
    for (int64_t i = 0; i < n; i++) out[i] = surf->sag(x[i], y[i]);
}

void ref_normal_n(const bt_surface* s, const double* x, const double* y,
                  double* out, int64_t n)
{
    Built b;
    const batoid::Surface* surf = build_surface(s, b);
    for (int64_t i = 0; i < n; i++)
        surf->normal(x[i], y[i], out[i], out[i+n], out[i+2*n]);
}

void ref_get_n_n(const bt_medium* m, const double* w, double* out, int64_t n)
{
    auto med = build_medium(m);
    for (int64_t i = 0; i < n; i++) out[i] = med->getN(w[i]);
}

void ref_contains_n(const bt_obsc* prog, int len, const double* x,
                    const double* y, uint8_t* out, int64_t n)
{
    ObscBuilt ob;
    const batoid::Obscuration* o = build_obsc(prog, len, ob);
    for (int64_t i = 0; i < n; i++) out[i] = o->contains(x[i], y[i]) ? 1 : 0;
}

void ref_apply_forward_transform(const double* dr, const double* drot,
                                 bt_rays rays, int nthreads)
{
    batoid::RayVector rv = make_rv(rays);
    batoid::applyForwardTransform(v3(dr), m3(drot), rv, nthreads);
}

void ref_apply_reverse_transform(const double* dr, const double* drot,
                                 bt_rays rays, int nthreads)
{
    batoid::RayVector rv = make_rv(rays);
    batoid::applyReverseTransform(v3(dr), m3(drot), rv, nthreads);
}

void ref_finish_parallel(const double* dr, const double* drot, const double* vv,
                         double* x, double* y, double* z, int64_t n)
{
    batoid::finishParallel(v3(dr), m3(drot), v3(vv), x, y, z, (size_t)n);
}

void ref_phase(bt_rays rays, double x, double y, double z, double t,
               double* out, int nthreads)
{
    batoid::RayVector rv = make_rv(rays);
    rv.phase(x, y, z, t, out, nthreads);
}

void ref_sum_amplitude(bt_rays rays, double x, double y, double z, double t,
                       int ignore_vignetted, double* out, int nthreads)
{
    batoid::RayVector rv = make_rv(rays);
    std::complex<double> a = rv.sumAmplitude(x, y, z, t, ignore_vignetted != 0, nthreads);
    out[0] = a.real();
    out[1] = a.imag();
}

}  // extern "C"
