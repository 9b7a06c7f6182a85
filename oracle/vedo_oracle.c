/*
 * vedo_oracle.c -- CPU FP64 restatement of VEDO's DEM step (TEST ORACLE).
 *
 * Test infrastructure: loaded only by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, as the checker.  Single-threaded like the
 * reference.  Build with -ffp-contract=off so that every expression rounds
 * exactly as the reference's x86-64 SSE2 build does.
 *
 * Reference quirks and the decisions taken here:
 *  - Vector3d::direction() of a zero vector is uninitialised in the reference
 *    (Vector3d.h:117-126); here it is the zero vector.
 *  - ISwModels is reached through an uninitialised pointer (ImpactSolver.h:70);
 *    its methods do not touch `this`, so they are free functions here.
 *  - The reference pairs IactPairTab[i] with cIact[i] when saving histories
 *    (NBSConsultant.cpp:384-390) although Initiate may skip pairs
 *    (SimMediator.cpp:686-706).  Here histories are saved from the created
 *    interactions themselves, which is what the reference does whenever no
 *    pair is skipped (always, for scenes whose every group pair has a model
 *    and no two density-0 DOs are paired).
 *  - NBS rmax is the range of the *last* local DO and its bounding-box max
 *    starts at numeric_limits<double>::min() (NBSConsultant.cpp:107-140);
 *    both are kept as written.
 */
#include "vedo_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* aries::math::_PI = 4*atan(1) (Constants.h:22) */

typedef struct { double x, y, z; } v3;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 ld(const double *p) { return mk(p[0], p[1], p[2]); }
static inline void st3(double *p, v3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
/* operator*(T s, Vector3d v) -> (s*v.x, ...) (Vector3d.h:313-316) */
static inline v3 smul(double s, v3 v) { return mk(s * v.x, s * v.y, s * v.z); }
/* Vector3d::operator*(T s) -> (v.x*s, ...) (Vector3d.h:268-271) */
static inline v3 muls(v3 v, double s) { return mk(v.x * s, v.y * s, v.z * s); }
static inline v3 neg(v3 v) { return mk(-v.x, -v.y, -v.z); }
/* Vector3d::dot (Vector3d.h:200-203) */
static inline double dot(v3 a, v3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
/* Vector3d::cross (Vector3d.h:215-218) */
static inline v3 cross(v3 a, v3 b)
{
    return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
/* Vector3d::length (Vector3d.h:112-115) */
static inline double len(v3 a) { return sqrt((a.x * a.x) + (a.y * a.y) + (a.z * a.z)); }
/* Vector3d::direction (Vector3d.h:117-126); zero for a zero vector */
static inline v3 dir(v3 a)
{
    double l = len(a);
    if (l != 0) return mk(a.x / l, a.y / l, a.z / l);
    return mk(0.0, 0.0, 0.0);
}
/* Vector3d::project_on (Vector3d.h:194-197) */
static inline v3 project_on(v3 a, v3 v) { return smul(dot(a, v) / len(v), dir(v)); }

/* Vector3d::rotate_around (Vector3d.h:152-163) */
static v3 rotate_around(v3 p, v3 va)
{
    v3 u = dir(va);
    double a = len(va);
    double s = cos(0.5 * a);
    double t = sin(0.5 * a);
    v3 v = muls(u, t);
    v3 t1 = smul(s * s, p);
    v3 t2 = smul(dot(p, v), v);
    v3 t3 = smul(2.0 * s, cross(v, p));
    v3 t4 = cross(v, cross(v, p));
    return add(add(add(t1, t2), t3), t4);
}

void vo_rotate_around(const double *p, const double *w, double *out)
{
    st3(out, rotate_around(ld(p), ld(w)));
}

/* ------------------------------------------------------------------------ */
/* Periodic boundary (SystemParameter PeriodicBoundaryConditions)           */
/* ------------------------------------------------------------------------ */

/* Boundary after Correct (Boundary.cpp:79-129): lower and upper point sorted
 * per axis, range = upper - lower, an axis periodic where they differ. */
typedef struct {
    int on[3], active;
    double lo[3], hi[3], len[3];
} pbc_t;

/* Boundary::DifferenceBoundaryConditions (Boundary.cpp:303-349): the
 * minimum image of a difference vector, one half range at a time. */
static double bc_diff1(double d, double l)
{
    if (d >= 0.5 * l) d -= l;
    if (d < -0.5 * l) d += l;
    return d;
}
static v3 bc_diff(const pbc_t *bc, v3 d)
{
    if (!bc) return d;
    if (bc->on[0]) d.x = bc_diff1(d.x, bc->len[0]);
    if (bc->on[1]) d.y = bc_diff1(d.y, bc->len[1]);
    if (bc->on[2]) d.z = bc_diff1(d.z, bc->len[2]);
    return d;
}

/* Boundary::EnforceBoundaryConditions (Boundary.cpp:225-301): a position
 * taken back into [lower, upper) by one range, or by fmod when further out. */
static double bc_wrap1(double x, double lo, double hi, double l)
{
    if (x >= hi) {
        x -= l;
        if (x >= hi) x = lo + fmod(x - lo, l);
    }
    if (x < lo) {
        x += l;
        if (x < lo) x = lo + fmod(x - lo, l) + l;
    }
    return x;
}
static v3 bc_enforce(const pbc_t *bc, v3 p)
{
    if (bc->on[0]) p.x = bc_wrap1(p.x, bc->lo[0], bc->hi[0], bc->len[0]);
    if (bc->on[1]) p.y = bc_wrap1(p.y, bc->lo[1], bc->hi[1], bc->len[1]);
    if (bc->on[2]) p.z = bc_wrap1(p.z, bc->lo[2], bc->hi[2], bc->len[2]);
    return p;
}

/* ------------------------------------------------------------------------ */
/* Contact detectors                                                        */
/* ------------------------------------------------------------------------ */

typedef struct {
    v3 n, p;
    double depth;
    int active;
} cinfo;

/* CDSphere_Sphere::CalDistance + Detect (CDSphere2Sphere.cpp:14-60); the
 * centre-to-centre vector is the minimum image under periodic boundaries
 * (:19-22, bc non-null once SimMediator::Initiate gave the interaction the
 * boundary, SimMediator.cpp:662-676). */
static void cd_ss(v3 xa, double ra, v3 xb, double rb, const pbc_t *bc, cinfo *c)
{
    v3 im = bc_diff(bc, sub(xb, xa));
    c->depth = ra + rb - len(im);
    c->n = dir(im);
    c->p = add(xa, smul(ra - 0.5 * c->depth, c->n));
    c->active = c->depth > 0;
}

/* CDSphere_QuasiPlate::CalDistance + Detect (CDSphere2QuasiPlate.cpp:15-97) */
static void cd_sp(v3 xa, double ra, v3 xb, v3 ox, v3 oz, double w, double h, double l, cinfo *c)
{
    v3 oy = cross(oz, ox);
    v3 cap = sub(xa, project_on(sub(xa, xb), oz));
    double hh = 0.5 * h, hw = 0.5 * w, hl = 0.5 * l;
    double dx = dot(sub(cap, xb), ox);
    if (dx < -hw) dx = -hw;
    else if (dx > hw) dx = hw;
    double dy = dot(sub(cap, xb), oy);
    if (dy < -hl) dy = -hl;
    else if (dy > hl) dy = hl;
    v3 caps = add(add(muls(ox, dx), muls(oy, dy)), xb);
    v3 im = sub(caps, xa);
    c->depth = ra + hh - len(im);
    c->n = dir(im);
    c->p = add(xa, smul(ra - 0.5 * c->depth, c->n));
    c->active = c->depth > 0;
}

/* CDSphere_QuasiPlateWithCircularHole::CalDistance + Detect
 * (CDSphere2QuasiPlateWithCircularHole.cpp:16-134).  As the plate, with the
 * closest point pushed to the hole's rim when the sphere's centre projects
 * into the hole; the hole's centre is xb + (hx, hy, 0) in global axes
 * (:46) while the rim point is placed along the plate's own axes. */
static void cd_sph(v3 xa, double ra, v3 xb, v3 ox, v3 oz, double w, double h, double l, double hr, double hx,
                   double hy, cinfo *c)
{
    v3 oy = cross(oz, ox);
    v3 cap = sub(xa, project_on(sub(xa, xb), oz));
    double hh = 0.5 * h, hw = 0.5 * w, hl = 0.5 * l;
    v3 ch = add(xb, mk(hx, hy, 0.0));
    v3 cha = sub(xa, ch);
    double dx = dot(sub(cap, xb), ox);
    if (dx < -hw) dx = -hw;
    else if (dx > hw) dx = hw;
    double dy = dot(sub(cap, xb), oy);
    if (dy < -hl) dy = -hl;
    else if (dy > hl) dy = hl;
    double dhx = dot(cha, ox), dhy = dot(cha, oy);
    double dh = sqrt(dhx * dhx + dhy * dhy);
    if (dh < hr) {
        if (dh == 0.0) {
            dx = hx;
            dy = hy;
        } else {
            dx = hx + dhx / dh * hr;
            dy = hy + dhy / dh * hr;
        }
    }
    v3 caps = add(add(muls(ox, dx), muls(oy, dy)), xb);
    v3 im = sub(caps, xa);
    if (dh == 0.0) c->depth = ra + hh - sqrt(len(im) * len(im) + hr * hr);
    else c->depth = ra + hh - len(im);
    c->n = dir(im);
    c->p = add(xa, smul(ra - 0.5 * c->depth, c->n));
    c->active = c->depth > 0;
}

/* CDSphere_QuasiCylinder::CalDistance + Detect (CDSphere2QuasiCylinder.cpp:16-80) */
static void cd_sc(v3 xa, double ra, v3 xb, v3 ax, double rc, double hc, cinfo *c)
{
    double hh = 0.5 * hc;
    double d = dot(sub(xa, xb), ax);
    if (d < -hh) d = -hh;
    else if (d > hh) d = hh;
    v3 caps = add(xb, smul(d, ax));
    v3 im = sub(caps, xa);
    c->depth = ra + rc - len(im);
    c->n = dir(im);
    c->p = add(xa, smul(ra - 0.5 * c->depth, c->n));
    c->active = c->depth > 0.0;
}

static void to_vo(const cinfo *ci, vo_contact *c)
{
    c->depth = ci->depth; c->active = ci->active;
    st3(c->n, ci->n); st3(c->p, ci->p);
}

void vo_cd_sphere_sphere(const double *xa, double ra, const double *xb, double rb, vo_contact *c)
{
    cinfo ci; cd_ss(ld(xa), ra, ld(xb), rb, NULL, &ci); to_vo(&ci, c);
}
void vo_cd_sphere_plate(const double *xa, double ra, const double *xb, const double *oxb,
                        const double *ozb, double w, double h, double l, vo_contact *c)
{
    cinfo ci; cd_sp(ld(xa), ra, ld(xb), ld(oxb), ld(ozb), w, h, l, &ci); to_vo(&ci, c);
}
void vo_cd_sphere_cylinder(const double *xa, double ra, const double *xb, const double *ozb,
                           double rc, double hc, vo_contact *c)
{
    cinfo ci; cd_sc(ld(xa), ra, ld(xb), ld(ozb), rc, hc, &ci); to_vo(&ci, c);
}

/* ------------------------------------------------------------------------ */
/* ImpactStatus helpers (ImpactStatus.cpp:6-9, 248-274)                     */
/* ------------------------------------------------------------------------ */

static void status_default(vo_status *s)
{
    memset(s, 0, sizeof(*s));
    s->last_contact = 1;   /* ImpactStatus() : bLastContact(true) */
    s->last_bond = 1;      /*                  bLastBond(true)    */
}

static void status_clean(vo_status *s)   /* ImpactStatus::Clean */
{
    memset(s, 0, sizeof(*s));
}

static void status_clean_contact(vo_status *s)   /* CleanContactStatusAndImpact */
{
    s->last_contact = s->contact;
    s->contact = 0;
    s->last_bond = s->bond;
    s->bond = 0;
    memset(s->shear, 0, sizeof(s->shear));
    memset(s->to_master, 0, sizeof(s->to_master));
    memset(s->ang_to_master, 0, sizeof(s->ang_to_master));
}

/* ------------------------------------------------------------------------ */
/* ISwModels helpers (ISwModels.cpp)                                         */
/* ------------------------------------------------------------------------ */

/* CriticalDamping(Ma, Mb, Kn, Ks) (ISwModels.cpp:88-108) */
static void critical_damping(double ma, double mb, double kn, double ks, double *cn, double *cs)
{
    double me;
    if (mb == 0) me = ma;
    else if (ma == 0) me = mb;
    else me = 0.5 * (ma + mb);
    *cn = 2.0 * sqrt(me * kn);
    *cs = 2.0 * sqrt(me * ks);
}

/* NormalBond (ISwModels.cpp:110-125) */
static int normal_bond(double kn, double bn, double depth)
{
    return (depth <= 0.0) && (kn != 0.0) && (bn != 0.0) && (-depth * kn < bn);
}

/* NormalForce 5-arg (ISwModels.cpp:134-139) */
static v3 normal_force(double kn, double cn, double depth, v3 n, v3 vn)
{
    return sub(smul(-kn * depth, n), smul(cn, vn));
}

/* NormalForce 6-arg (ISwModels.cpp:163-176) */
static v3 normal_force_bond(double kn, double cn, double depth, v3 n, v3 vn)
{
    if (dot(vn, n) < 0.0) return sub(smul(-kn * depth, n), smul(cn, vn));
    return smul(-cn, vn);
}

/* NormalForceHertzSpring (ISwModels.cpp:127-132) */
static v3 normal_force_hertz(double kn, double cn, double depth, v3 n, v3 vn)
{
    return sub(smul(-kn * pow(depth, 1.5), n), smul(cn, vn));
}

/* ShearForceRotation (ISwModels.cpp:263-273) */
static v3 shear_rotation(v3 f, v3 d)
{
    if (dot(f, d) <= 0.0) return smul(-len(f), dir(d));
    return smul(len(f), dir(d));
}

/* ------------------------------------------------------------------------ */
/* Impact solvers                                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    v3 imp_a, ang_a, ang_b;   /* impulse on A; angular impulses on A and B */
    int added;                /* AddImpact called on A and B */
    int linear_only;          /* ISwHertz adds no angular impulse */
} solve_out;

typedef struct { v3 x, v, w; double m; } body;

/* ISwBSDBF / ISwBSDB2F / ISwBtSDBF ::NextStep
 * (ISwBSDBF.cpp:26-139, ISwBSDB2F.cpp:26-149, ISwBtSDBF.cpp:26-137) */
static void solve_sd(int law, const double *p, const body *A, const body *B, const cinfo *ci,
                     double dt, const pbc_t *bc, vo_status *st, solve_out *o)
{
    const double kn = p[0], ks = p[1], bn = p[4], bs = p[5], fc = p[6];
    double cn, cs;
    critical_damping(A->m, B->m, kn, ks, &cn, &cs);   /* ISwBSDBF.cpp:12-15 */
    cn = p[2] * cn;
    cs = p[3] * cs;
    const double depth = ci->depth;
    const v3 n = ci->n;
    o->added = 0; o->linear_only = 0;
    if (!ci->active && !normal_bond(kn, bn, depth)) {
        status_clean(st);
        o->imp_a = o->ang_a = o->ang_b = mk(0, 0, 0);
        return;
    }
    st3(st->point, ci->p); st3(st->dir, ci->n); st->overlap = ci->depth;
    v3 ra = mk(0, 0, 0), rb = mk(0, 0, 0), fn = mk(0, 0, 0), fs = mk(0, 0, 0);
    if (depth > 0.0) {
        /* lever arms as minimum images under periodic boundaries (ISwBSDBF.cpp:69-73) */
        ra = bc_diff(bc, sub(ci->p, A->x));
        rb = bc_diff(bc, sub(ci->p, B->x));
        v3 rv = sub(add(sub(A->v, B->v), cross(A->w, ra)), cross(B->w, rb));
        v3 vn = smul(dot(rv, n), n);
        v3 vs = sub(rv, vn);
        fn = normal_force(kn, cn, depth, n, vn);
        if (law == VO_BTSDBF)
            fs = sub(smul(-ks * dt, vs), smul(cs, vs));
        else
            fs = sub(shear_rotation(ld(st->shear), vs), smul(ks * dt, vs));
        double fmax = fc * kn * depth;
        if (bs == 0.0) {
            if (len(fs) > fmax) fs = smul(fmax, dir(fs));
            else fs = sub(fs, smul(cs, vs));
        } else {
            if (len(fs) > bs) fs = smul(fmax, dir(fs));
            else fs = sub(fs, smul(cs, vs));
        }
        if (law != VO_BTSDBF) st3(st->shear, fs);
        st->contact = 1;
    } else if (law != VO_BSDB2F || st->last_bond) {
        v3 rv = sub(A->v, B->v);
        v3 vn = smul(dot(rv, n), n);
        if (law == VO_BTSDBF) fn = normal_force(kn, cn, depth, n, vn);
        else fn = normal_force_bond(kn, cn, depth, n, vn);
        fs = mk(0, 0, 0);
        status_clean_contact(st);
    }
    v3 f = add(fn, fs);
    v3 imp = muls(f, dt);
    o->imp_a = imp;
    o->ang_a = smul(dt, cross(ra, f));
    o->ang_b = smul(-dt, cross(rb, f));
    st3(st->to_master, smul(-1.0, imp));
    st3(st->ang_to_master, o->ang_b);
    o->added = 1;
}

/* ISwH::NextStep (ISwH.cpp:30-156) */
static void solve_h(const double *p, const body *A, const body *B, const cinfo *ci, double dt,
                    const pbc_t *bc, vo_status *st, solve_out *o)
{
    const double knh = p[0], meff = p[1], ksc1 = p[2], ksc2 = p[3], betan = p[4], fk = p[5];
    const double depth = ci->depth;
    const v3 n = ci->n;
    o->added = 0; o->linear_only = 0;
    if (!ci->active) {
        status_clean(st);
        o->imp_a = o->ang_a = o->ang_b = mk(0, 0, 0);
        return;
    }
    st3(st->point, ci->p); st3(st->dir, ci->n); st->overlap = ci->depth;
    v3 ra = mk(0, 0, 0), rb = mk(0, 0, 0), fn = mk(0, 0, 0), fs = mk(0, 0, 0);
    if (depth > 0.0) {
        ra = bc_diff(bc, sub(ci->p, A->x));   /* ISwH.cpp:69-73 */
        rb = bc_diff(bc, sub(ci->p, B->x));
        v3 rv = sub(add(sub(A->v, B->v), cross(A->w, ra)), cross(B->w, rb));
        v3 vn = smul(dot(rv, n), n);
        v3 vs = sub(rv, vn);
        double kn;
        if (!st->contact) {
            if (len(vn) == 0.0) kn = knh;
            else if (betan < 0.99)
                kn = 0.957832746294131 / (1.0 - betan * betan) * pow(meff, 0.2) * pow(knh, 0.8) *
                     pow(len(vn), 0.4);
            else
                kn = 479.155951122633 * pow(meff, 0.2) * pow(knh, 0.8) * pow(len(vn), 0.4);
            st->kn = kn;
            st->contact = 1;
        } else {
            kn = st->kn;
        }
        double ncd = 2.0 * sqrt(meff * kn);
        double ks = ksc1 * pow(ksc2 * fabs(kn * depth), 1.0 / 3.0);
        fn = normal_force(kn, ncd * betan, depth, n, vn);
        double ff = fk * kn * depth;
        fs = sub(shear_rotation(ld(st->shear), vs), smul(ks * dt, vs));
        if (len(fs) >= ff) fs = smul(ff, dir(fs));
        st3(st->shear, fs);
    } else {
        status_clean_contact(st);
    }
    v3 f = add(fn, fs);
    v3 imp = muls(f, dt);
    o->imp_a = imp;
    o->ang_a = smul(dt, cross(ra, f));
    o->ang_b = smul(-dt, cross(rb, f));
    st3(st->to_master, smul(-1.0, imp));
    st3(st->ang_to_master, o->ang_b);
    o->added = 1;
}

/* ISwHertz::NextStep (ISwHertz.cpp:20-94) */
static void solve_hertz(const double *p, const body *A, const body *B, const cinfo *ci, double dt,
                        const pbc_t *bc, vo_status *st, solve_out *o)
{
    const double kn = p[0], cn = p[1];
    const double depth = ci->depth;
    const v3 n = ci->n;
    o->added = 0; o->linear_only = 1;
    o->ang_a = o->ang_b = mk(0, 0, 0);
    if (!ci->active) {
        status_clean(st);
        o->imp_a = mk(0, 0, 0);
        return;
    }
    st3(st->point, ci->p); st3(st->dir, ci->n); st->overlap = ci->depth;
    v3 fn = mk(0, 0, 0);
    if (depth > 0.0) {
        v3 ra = bc_diff(bc, sub(ci->p, A->x));   /* ISwHertz.cpp:59-63 */
        v3 rb = bc_diff(bc, sub(ci->p, B->x));
        v3 rv = sub(add(sub(A->v, B->v), cross(A->w, ra)), cross(B->w, rb));
        v3 vn = smul(dot(rv, n), n);
        if (!st->contact) {
            st->contact = 1;
            st->init_vel = len(vn);
        }
        fn = normal_force_hertz(kn, cn, depth, n, vn);
    } else {
        status_clean_contact(st);
    }
    v3 imp = muls(fn, dt);
    o->imp_a = imp;
    st3(st->to_master, neg(imp));
    memset(st->ang_to_master, 0, sizeof(st->ang_to_master));
    o->added = 1;
}

static void solve(int law, const double *p, const body *A, const body *B, const cinfo *ci,
                  double dt, const pbc_t *bc, vo_status *st, solve_out *o)
{
    switch (law) {
    case VO_BSDBF: case VO_BSDB2F: case VO_BTSDBF: solve_sd(law, p, A, B, ci, dt, bc, st, o); break;
    case VO_H: solve_h(p, A, B, ci, dt, bc, st, o); break;
    case VO_HERTZ: solve_hertz(p, A, B, ci, dt, bc, st, o); break;
    default: memset(o, 0, sizeof(*o)); break;
    }
}

int vo_solve(int law, const double *p, double ma, double mb, const vo_contact *c,
             const double *xa, const double *va, const double *wa,
             const double *xb, const double *vb, const double *wb,
             double dt, vo_status *st, double *imp_a, double *ang_a, double *ang_b)
{
    body A = {ld(xa), ld(va), ld(wa), ma}, B = {ld(xb), ld(vb), ld(wb), mb};
    cinfo ci; ci.depth = c->depth; ci.n = ld(c->n); ci.p = ld(c->p); ci.active = c->active;
    solve_out o;
    solve(law, p, &A, &B, &ci, dt, NULL, st, &o);
    st3(imp_a, o.imp_a); st3(ang_a, o.ang_a); st3(ang_b, o.ang_b);
    return o.added;
}

/* ------------------------------------------------------------------------ */
/* Simulation state                                                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    unsigned slave, master;   /* DO ids (Assembler::CreateInteraction ordering) */
    unsigned lo, hi;          /* IactPairTab key */
    int law;
    const double *p;
    cinfo ci;
    vo_status st;
} iact;

typedef struct {
    unsigned lo, hi;
    int order;   /* position among the keys ModifyPair moves (ties: the later wins) */
    vo_status st;
} record;

struct vo_sim {
    int n_do, n_models, n_groups;
    vo_model *models;
    vo_iact *iact;
    int *model_of;
    /* derived model quantities (DOModel::Analysis, DOModel.cpp:17-149) */
    double *range, *mass, *sudo;
    /* per-DO state and accumulators (DOStatus + DiscreteObject) */
    v3 *pos, *vel, *ox, *oz, *av, *imp, *aimp, *last_imp, *last_aimp;
    double *do_sudo;          /* DiscreteObject::dSudoMass (0 for fixed/orbital) */
    double dt, t_start, t_stop, t_current;
    v3 g;
    double sf;
    unsigned up_iact, record_every, round;
    int consultant;
    /* candidates (IactPairTab) and interactions (cIact) */
    unsigned *cand_lo, *cand_hi;
    int n_cand, cap_cand;
    iact *ia;
    int n_ia;
    /* IactRecordTab */
    record *rec;
    int n_rec;
    /* zone of interest (Boundary, Boundary.cpp:79-129) and the scene's DO id of each current DO */
    int zoi_on[3], zoi_active;
    double zoi_lo[3], zoi_hi[3];
    /* periodic boundary; interactions see it only while it is active */
    pbc_t pbc;
    unsigned *orig_id;
    /* CleanUp counts: DOs erased, records moved, moved records of an erased DO,
       kept records replaced by a moved one */
    long long cu_erased, cu_moved, cu_from_dead, cu_replaced;
};

static const pbc_t *sim_bc(const vo_sim *s) { return s->pbc.active ? &s->pbc : NULL; }

static double model_range(const vo_model *m)
{
    switch (m->shape) {
    case VO_SPHERE: return m->attr[0];
    case VO_QUASICYLINDER:
        if (m->scope == VO_LOCAL) return fmax(m->attr[0], 0.5 * m->attr[1]);
        return m->attr[0];
    case VO_QUASIPLATE:
    case VO_QUASIPLATE_HOLE:   /* DOModel.cpp:90-106 */
        if (m->scope == VO_LOCAL) {
            double a = m->attr[0] > m->attr[1] ? m->attr[0] : m->attr[1];
            a = a > m->attr[2] ? a : m->attr[2];
            return 0.5 * a;
        }
        return 0.5 * m->attr[1];
    default: return 0.0;
    }
}

static void model_mass(const vo_model *m, double *mass, double *sudo)
{
    const double PI = 4.0 * atan(1.0);
    double vol = 0.0, r, h, w, l;
    switch (m->shape) {
    case VO_SPHERE:
        r = m->attr[0];
        vol = (4.0 / 3.0 * PI) * r * r * r;   /* _FourthThirdsPI * r*r*r */
        break;
    case VO_QUASICYLINDER:
        r = m->attr[0]; h = m->attr[1];
        vol = PI * r * r * h;
        break;
    case VO_QUASIPLATE:
        w = m->attr[0]; h = m->attr[1]; l = m->attr[2];
        vol = w * h * l;
        break;
    case VO_QUASIPLATE_HOLE:   /* DOModel.cpp:108-116 */
        w = m->attr[0]; h = m->attr[1]; l = m->attr[2];
        r = m->hole[0];
        if (fabs(m->hole[1]) <= (0.5 * w - r) && fabs(m->hole[2]) <= (0.5 * l - r)) vol = w * h * l - r * r * PI * h;
        else vol = w * h * l;
        break;
    }
    *mass = m->density * vol;
    *sudo = m->density_factor * *mass;
}

static const vo_iact *iact_of(const vo_sim *s, unsigned a, unsigned b)
{
    int ga = s->models[s->model_of[a]].group, gb = s->models[s->model_of[b]].group;
    return &s->iact[ga * s->n_groups + gb];
}

static int push_cand(vo_sim *s, unsigned a, unsigned b)
{
    if (s->n_cand == s->cap_cand) {
        int nc = s->cap_cand ? 2 * s->cap_cand : 1024;
        unsigned *lo = realloc(s->cand_lo, sizeof(unsigned) * nc);
        unsigned *hi = realloc(s->cand_hi, sizeof(unsigned) * nc);
        if (!lo || !hi) { free(lo ? lo : NULL); return -1; }
        s->cand_lo = lo; s->cand_hi = hi; s->cap_cand = nc;
    }
    s->cand_lo[s->n_cand] = a < b ? a : b;
    s->cand_hi[s->n_cand] = a < b ? b : a;
    s->n_cand++;
    return 0;
}

/* DOMap::CalDistance (DOMap.cpp:142-414); with a boundary (the Leap
 * consultant passes the periodic one, LeapConsultant.cpp:199-212; NBS none,
 * NBSConsultant.cpp:330) every distance vector is its minimum image */
static double domap_distance(const vo_sim *s, unsigned a, unsigned b, const pbc_t *bc)
{
    const vo_model *ma = &s->models[s->model_of[a]], *mb = &s->models[s->model_of[b]];
    if (ma->shape > mb->shape) { unsigned t = a; a = b; b = t; const vo_model *tm = ma; ma = mb; mb = tm; }
    if (ma->shape == VO_SPHERE && mb->shape == VO_SPHERE) {
        return len(bc_diff(bc, sub(s->pos[a], s->pos[b])));
    }
    /* the plate with a hole: the plate's distance, the hole ignored (DOMap.cpp:336-374) */
    if (ma->shape == VO_SPHERE && (mb->shape == VO_QUASIPLATE || mb->shape == VO_QUASIPLATE_HOLE)) {
        v3 ca = s->pos[a], cb = s->pos[b], oz = s->oz[b], ox = s->ox[b];
        v3 oy = cross(oz, ox);
        v3 cap = sub(ca, project_on(sub(ca, cb), oz));
        double hw = 0.5 * mb->attr[0], hl = 0.5 * mb->attr[2];
        double dx = dot(sub(cap, cb), ox);
        if (dx < -hw) dx = -hw; else if (dx > hw) dx = hw;
        double dy = dot(sub(cap, cb), oy);
        if (dy < -hl) dy = -hl; else if (dy > hl) dy = hl;
        v3 caps = add(add(muls(ox, dx), muls(oy, dy)), cb);
        return len(bc_diff(bc, sub(caps, ca)));
    }
    if (ma->shape == VO_SPHERE && mb->shape == VO_QUASICYLINDER) {
        double hh = 0.5 * mb->attr[1];
        v3 ca = s->pos[a], cb = s->pos[b], ax = s->oz[b];
        double d = dot(sub(ca, cb), ax);
        v3 cap = add(cb, muls(ax, d));
        v3 im;
        if (d < hh && d > -hh) im = sub(cap, ca);
        else if (d >= hh) im = sub(add(cb, muls(ax, hh)), ca);
        else im = sub(sub(cb, muls(ax, hh)), ca);
        return len(bc_diff(bc, im));
    }
    return INFINITY;
}

/* BuildIactTab(v1, v2) and BuildIactTab(v) (LeapConsultant.cpp:198-262,
 * NBSConsultant.cpp:318-374) */
static int build_pairs(vo_sim *s, const unsigned *v1, int n1, const unsigned *v2, int n2,
                       const double *safe, const pbc_t *bc)
{
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j) {
            unsigned a = v1[i], b = v2[j];
            if (iact_of(s, a, b)->law == VO_LAW_NONE) continue;
            if (domap_distance(s, a, b, bc) > safe[a] + safe[b]) continue;
            if (push_cand(s, a, b)) return -1;
        }
    return 0;
}
static int build_pairs_self(vo_sim *s, const unsigned *v, int n, const double *safe, const pbc_t *bc)
{
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            unsigned a = v[i], b = v[j];
            if (iact_of(s, a, b)->law == VO_LAW_NONE) continue;
            if (domap_distance(s, a, b, bc) > safe[a] + safe[b]) continue;
            if (push_cand(s, a, b)) return -1;
        }
    return 0;
}

typedef struct { unsigned k[3]; unsigned id; } cellkey;   /* k in comparator order */

static int cmp_cell(const void *pa, const void *pb)
{
    const cellkey *a = pa, *b = pb;
    for (int d = 0; d < 3; ++d)
        if (a->k[d] != b->k[d]) return a->k[d] < b->k[d] ? -1 : 1;
    return a->id < b->id ? -1 : (a->id > b->id);
}

static int cmp_double(const void *pa, const void *pb)
{
    double a = *(const double *)pa, b = *(const double *)pb;
    return (a > b) - (a < b);
}

/* Binary search for the cell (ix,iy,iz) in the sorted key array; returns the
 * first index or -1.  perm maps (x,y,z) to comparator order. */
static int find_cell(const cellkey *ck, int n, const int *perm, unsigned x, unsigned y, unsigned z)
{
    unsigned xyz[3] = {x, y, z};
    unsigned key[3] = {xyz[perm[0]], xyz[perm[1]], xyz[perm[2]]};
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        int c = 0;
        for (int d = 0; d < 3 && !c; ++d)
            if (ck[mid].k[d] != key[d]) c = ck[mid].k[d] < key[d] ? -1 : 1;
        if (c < 0) lo = mid + 1; else hi = mid;
    }
    if (lo < n && ck[lo].k[0] == key[0] && ck[lo].k[1] == key[1] && ck[lo].k[2] == key[2]) return lo;
    return -1;
}

/* Shared grid traversal. offsets: 13 neighbour (dx,dy,dz) in reference order;
 * self/global inserted at position `self_at`. */
typedef struct {
    cellkey *ck; int nck;
    int *cell_start, *cell_len; int ncell;   /* per distinct cell (sorted) */
    unsigned *ids;                          /* DO ids grouped by cell */
    int perm[3];
} grid;

static void grid_free(grid *g) { free(g->ck); free(g->cell_start); free(g->cell_len); free(g->ids); }

static int grid_build(const vo_sim *s, grid *g, const unsigned *loc, int nloc, double xmin, double ymin,
                      double zmin, double zr, const int perm[3])
{
    memset(g, 0, sizeof(*g));
    g->perm[0] = perm[0]; g->perm[1] = perm[1]; g->perm[2] = perm[2];
    g->ck = malloc(sizeof(cellkey) * (nloc ? nloc : 1));
    g->ids = malloc(sizeof(unsigned) * (nloc ? nloc : 1));
    g->cell_start = malloc(sizeof(int) * (nloc ? nloc : 1));
    g->cell_len = malloc(sizeof(int) * (nloc ? nloc : 1));
    if (!g->ck || !g->ids || !g->cell_start || !g->cell_len) return -1;
    for (int i = 0; i < nloc; ++i) {
        v3 p = s->pos[loc[i]];
        unsigned xyz[3] = {(unsigned)((p.x - xmin) / zr), (unsigned)((p.y - ymin) / zr),
                           (unsigned)((p.z - zmin) / zr)};
        g->ck[i].k[0] = xyz[perm[0]]; g->ck[i].k[1] = xyz[perm[1]]; g->ck[i].k[2] = xyz[perm[2]];
        g->ck[i].id = loc[i];
    }
    qsort(g->ck, nloc, sizeof(cellkey), cmp_cell);
    int nc = 0;
    for (int i = 0; i < nloc; ++i) {
        g->ids[i] = g->ck[i].id;
        if (i == 0 || memcmp(g->ck[i].k, g->ck[i - 1].k, sizeof(g->ck[i].k))) {
            g->cell_start[nc] = i; g->cell_len[nc] = 0;
            g->ck[nc] = g->ck[i];   /* compact distinct keys in front (i >= nc) */
            nc++;
        }
        g->cell_len[nc - 1]++;
    }
    g->ncell = nc; g->nck = nc;
    return 0;
}

static void unperm(const grid *g, int c, unsigned *x, unsigned *y, unsigned *z)
{
    unsigned xyz[3];
    for (int d = 0; d < 3; ++d) xyz[g->perm[d]] = g->ck[c].k[d];
    *x = xyz[0]; *y = xyz[1]; *z = xyz[2];
}

static const int NB13[13][3] = {
    {-1, -1, -1}, {-1, -1, 0}, {-1, -1, 1}, {0, -1, -1}, {0, -1, 0}, {0, -1, 1},
    {1, -1, -1}, {1, -1, 0}, {1, -1, 1}, {-1, 0, -1}, {-1, 0, 0}, {0, 0, -1}, {1, 0, -1}};

/* safe distance per DO: NBSConsultant.cpp:133-135 / LeapConsultant.h:117-119 */
static void safe_lengths(const vo_sim *s, double *safe)
{
    double dtu = s->up_iact * s->dt;
    double glen = len(s->g);
    for (int i = 0; i < s->n_do; ++i) {
        double rng = s->range[s->model_of[i]];
        safe[i] = s->sf * (2.0 * rng + dtu * len(s->vel[i]) + 0.5 * dtu * dtu * glen);
    }
}

/* LeapConsultant::Reset + subReset (LeapConsultant.cpp:108-195, LeapConsultant.h:57-514).
 * Periodic boundary: an axis takes part where it is periodic and the spheres
 * span at least half its range (LeapConsultant.h:146-150); a half-shell cell
 * that does not exist is replaced, for a cell on a border the offset crosses,
 * by the cell whose index on every periodic offset axis is the far end
 * (max for -1, 0 for +1) and on the other offset axes the plain neighbour,
 * unless that index is the cell's other neighbour on some offset axis
 * (LeapConsultant.h:199-444).  The far-end index is taken on every periodic
 * offset axis, also where the cell is not on that axis's border. */
static int reset_leap(vo_sim *s)
{
    s->n_cand = 0;
    int n = s->n_do, nloc = 0, nglob = 0;
    unsigned *loc = malloc(sizeof(unsigned) * (n ? n : 1));
    unsigned *glob = malloc(sizeof(unsigned) * (n ? n : 1));
    double *safe = malloc(sizeof(double) * (n ? n : 1));
    double *vx = malloc(sizeof(double) * (n ? n : 1)), *vy = malloc(sizeof(double) * (n ? n : 1));
    double *vz = malloc(sizeof(double) * (n ? n : 1)), *vv = malloc(sizeof(double) * (n ? n : 1));
    double *vr = malloc(sizeof(double) * (n ? n : 1));
    int rc = -1;
    grid g; memset(&g, 0, sizeof(g));
    if (!loc || !glob || !safe || !vx || !vy || !vz || !vv || !vr) goto out;
    for (int i = 0; i < n; ++i) {
        const vo_model *m = &s->models[s->model_of[i]];
        if (m->scope == VO_LOCAL) {
            vx[nloc] = s->pos[i].x; vy[nloc] = s->pos[i].y; vz[nloc] = s->pos[i].z;
            vv[nloc] = len(s->vel[i]); vr[nloc] = s->range[s->model_of[i]];
            loc[nloc++] = (unsigned)i;
        } else {
            glob[nglob++] = (unsigned)i;
        }
    }
    if (nloc == 0) { rc = 0; goto out; }   /* the reference indexes an empty vector here */
    qsort(vx, nloc, sizeof(double), cmp_double); qsort(vy, nloc, sizeof(double), cmp_double);
    qsort(vz, nloc, sizeof(double), cmp_double); qsort(vv, nloc, sizeof(double), cmp_double);
    qsort(vr, nloc, sizeof(double), cmp_double);
    double xmin = vx[0], xmax = vx[nloc - 1], ymin = vy[0], ymax = vy[nloc - 1];
    double zmin = vz[0], zmax = vz[nloc - 1], vmax = vv[nloc - 1], rmax = vr[nloc - 1];
    const double discard = 0.02, keep = 1.0 - discard;
    size_t ihi = (size_t)(keep * nloc - 1), ilo = (size_t)(discard * nloc);
    double xs = vx[ihi] - vx[ilo], ys = vy[ihi] - vy[ilo], zs = vz[ihi] - vz[ilo];
    int perm[3];
    /* TrirLeX: x,y,z  TrirLeY: y,x,z  TrirLeZ: z,x,y (NBSConsultant.h:52-118) */
    if (xs >= ys) {
        if (xs >= zs) { perm[0] = 0; perm[1] = 1; perm[2] = 2; }
        else { perm[0] = 2; perm[1] = 0; perm[2] = 1; }
    } else {
        if (ys >= zs) { perm[0] = 1; perm[1] = 0; perm[2] = 2; }
        else { perm[0] = 2; perm[1] = 0; perm[2] = 1; }
    }
    safe_lengths(s, safe);
    double dtu = s->up_iact * s->dt;
    double zr = s->sf * (2.0 * rmax + dtu * vmax + 0.5 * dtu * dtu * len(s->g));
    const pbc_t *bc = sim_bc(s);
    /* periodic axes of the traversal and the last zone index per axis */
    const double span[3] = {xmax - xmin, ymax - ymin, zmax - zmin};
    int pax[3], has_pbc = 0;
    unsigned maxid[3];
    for (int d = 0; d < 3; ++d) {
        pax[d] = s->pbc.on[d] && span[d] >= 0.5 * s->pbc.len[d];
        has_pbc = has_pbc || pax[d];
        unsigned nc = (unsigned)ceil(span[d] / zr);
        maxid[d] = (nc > 1u ? nc : 1u) - 1u;
    }
    if (grid_build(s, &g, loc, nloc, xmin, ymin, zmin, zr, perm)) goto out;
    for (int c = 0; c < g.ncell; ++c) {
        unsigned ix, iy, iz;
        unperm(&g, c, &ix, &iy, &iz);
        const unsigned id[3] = {ix, iy, iz};
        const unsigned *v1 = g.ids + g.cell_start[c];
        int n1 = g.cell_len[c];
        for (int k = 0; k < 13; ++k) {
            int c2 = find_cell(g.ck, g.ncell, g.perm, ix + NB13[k][0], iy + NB13[k][1], iz + NB13[k][2]);
            if (c2 < 0 && has_pbc) {
                int border = 0, dup = 0;
                unsigned w[3];
                for (int d = 0; d < 3; ++d) {
                    const int o = NB13[k][d];
                    if (o < 0) {
                        border = border || id[d] == 0u;
                        w[d] = pax[d] ? maxid[d] : id[d] - 1u;
                        dup = dup || w[d] == id[d] + 1u;
                    } else if (o > 0) {
                        border = border || id[d] == maxid[d];
                        w[d] = pax[d] ? 0u : id[d] + 1u;
                        dup = dup || w[d] == id[d] - 1u;
                    } else {
                        w[d] = id[d];
                    }
                }
                if (border && !dup) c2 = find_cell(g.ck, g.ncell, g.perm, w[0], w[1], w[2]);
            }
            if (c2 >= 0 && build_pairs(s, v1, n1, g.ids + g.cell_start[c2], g.cell_len[c2], safe, bc)) goto out;
        }
        if (build_pairs_self(s, v1, n1, safe, bc)) goto out;
        if (build_pairs(s, v1, n1, glob, nglob, safe, bc)) goto out;
    }
    rc = 0;
out:
    grid_free(&g);
    free(loc); free(glob); free(safe); free(vx); free(vy); free(vz); free(vv); free(vr);
    return rc;
}

/* NBSConsultant::Reset (NBSConsultant.cpp:85-316) */
static int reset_nbs(vo_sim *s)
{
    s->n_cand = 0;
    int n = s->n_do, nloc = 0, nglob = 0;
    unsigned *loc = malloc(sizeof(unsigned) * (n ? n : 1));
    unsigned *glob = malloc(sizeof(unsigned) * (n ? n : 1));
    double *safe = malloc(sizeof(double) * (n ? n : 1));
    int rc = -1;
    grid g; memset(&g, 0, sizeof(g));
    if (!loc || !glob || !safe) goto out;
    double xmin = DBL_MAX, xmax = DBL_MIN, ymin = DBL_MAX, ymax = DBL_MIN, zmin = DBL_MAX, zmax = DBL_MIN;
    double vmax = 0.0, rmax = 0.0;
    for (int i = 0; i < n; ++i) {
        const vo_model *m = &s->models[s->model_of[i]];
        if (m->scope == VO_LOCAL) {
            v3 p = s->pos[i];
            xmin = fmin(xmin, p.x); xmax = fmax(xmax, p.x);
            ymin = fmin(ymin, p.y); ymax = fmax(ymax, p.y);
            zmin = fmin(zmin, p.z); zmax = fmax(zmax, p.z);
            double vl = len(s->vel[i]);
            vmax = vmax > vl ? vmax : vl;
            rmax = s->range[s->model_of[i]];
            loc[nloc++] = (unsigned)i;
        } else {
            glob[nglob++] = (unsigned)i;
        }
    }
    safe_lengths(s, safe);
    double dtu = s->up_iact * s->dt;
    double zr = s->sf * (2.0 * rmax + dtu * vmax + 0.5 * dtu * dtu * len(s->g));
    unsigned ncx = (unsigned)ceil((xmax - xmin) / zr); if (ncx < 1) ncx = 1;
    unsigned ncy = (unsigned)ceil((ymax - ymin) / zr); if (ncy < 1) ncy = 1;
    unsigned ncz = (unsigned)ceil((zmax - zmin) / zr); if (ncz < 1) ncz = 1;
    int perm[3] = {1, 0, 2};   /* TrirLeY; only affects lookup, not traversal */
    if (grid_build(s, &g, loc, nloc, xmin, ymin, zmin, zr, perm)) goto out;
    for (unsigned ix = 0; ix < ncx; ++ix)
        for (unsigned iy = 0; iy < ncy; ++iy)
            for (unsigned iz = 0; iz < ncz; ++iz) {
                int c = find_cell(g.ck, g.ncell, g.perm, ix, iy, iz);
                if (c < 0) continue;
                const unsigned *v1 = g.ids + g.cell_start[c];
                int n1 = g.cell_len[c];
                for (int k = 0; k < 12; ++k) {
                    int c2 = find_cell(g.ck, g.ncell, g.perm, ix + NB13[k][0], iy + NB13[k][1], iz + NB13[k][2]);
                    if (c2 >= 0 && build_pairs(s, v1, n1, g.ids + g.cell_start[c2], g.cell_len[c2], safe, NULL)) goto out;
                }
                if (build_pairs_self(s, v1, n1, safe, NULL)) goto out;
                if (build_pairs(s, v1, n1, glob, nglob, safe, NULL)) goto out;
                int c2 = find_cell(g.ck, g.ncell, g.perm, ix + 1, iy, iz - 1);
                if (c2 >= 0 && build_pairs(s, v1, n1, g.ids + g.cell_start[c2], g.cell_len[c2], safe, NULL)) goto out;
            }
    rc = 0;
out:
    grid_free(&g);
    free(loc); free(glob); free(safe);
    return rc;
}

static int cmp_rec(const void *pa, const void *pb)
{
    const record *a = pa, *b = pb;
    if (a->lo != b->lo) return a->lo < b->lo ? -1 : 1;
    return (a->hi > b->hi) - (a->hi < b->hi);
}

static int cmp_rec_first(const void *pa, const void *pb)
{
    const int c = cmp_rec(pa, pb);
    if (c) return c;
    const record *a = pa, *b = pb;
    return (a->order > b->order) - (a->order < b->order);
}

static const record *find_rec(const vo_sim *s, unsigned lo, unsigned hi)
{
    int a = 0, b = s->n_rec;
    while (a < b) {
        int m = (a + b) / 2;
        if (s->rec[m].lo < lo || (s->rec[m].lo == lo && s->rec[m].hi < hi)) a = m + 1; else b = m;
    }
    if (a < s->n_rec && s->rec[a].lo == lo && s->rec[a].hi == hi) return &s->rec[a];
    return NULL;
}

static void detect(vo_sim *s, iact *it)
{
    unsigned a = it->slave, b = it->master;
    const vo_model *ma = &s->models[s->model_of[a]], *mb = &s->models[s->model_of[b]];
    if (mb->shape == VO_SPHERE)
        cd_ss(s->pos[a], ma->attr[0], s->pos[b], mb->attr[0], sim_bc(s), &it->ci);
    else if (mb->shape == VO_QUASIPLATE)
        cd_sp(s->pos[a], ma->attr[0], s->pos[b], s->ox[b], s->oz[b], mb->attr[0], mb->attr[1], mb->attr[2], &it->ci);
    else if (mb->shape == VO_QUASICYLINDER)
        cd_sc(s->pos[a], ma->attr[0], s->pos[b], s->oz[b], mb->attr[0], mb->attr[1], &it->ci);
    else if (mb->shape == VO_QUASIPLATE_HOLE)
        cd_sph(s->pos[a], ma->attr[0], s->pos[b], s->ox[b], s->oz[b], mb->attr[0], mb->attr[1], mb->attr[2],
               mb->hole[0], mb->hole[1], mb->hole[2], &it->ci);
}

/* SimMediator::Initiate (SimMediator.cpp:640-712) + Assembler::CreateInteraction
 * (Assembler.cpp:15-87). */
static int initiate(vo_sim *s)
{
    free(s->ia);
    s->ia = malloc(sizeof(iact) * (s->n_cand ? s->n_cand : 1));
    if (!s->ia) return -1;
    s->n_ia = 0;
    for (int k = 0; k < s->n_cand; ++k) {
        unsigned p1 = s->cand_lo[k], p2 = s->cand_hi[k];
        const vo_iact *im = iact_of(s, p1, p2);
        if (im->law == VO_LAW_NONE) continue;
        const vo_model *m1 = &s->models[s->model_of[p1]], *m2 = &s->models[s->model_of[p2]];
        if (p1 == p2) continue;
        if (m2->density < 0.0 || m1->density < 0.0) continue;
        if (m2->density == 0.0 && m1->density == 0.0) continue;
        unsigned sl, ms;
        if (m1->shape > m2->shape) { sl = p2; ms = p1; }
        else if (m1->shape < m2->shape) { sl = p1; ms = p2; }
        else if (m1->density == 0.0) { sl = p2; ms = p1; }
        else { sl = p1; ms = p2; }
        const vo_model *msl = &s->models[s->model_of[sl]], *mms = &s->models[s->model_of[ms]];
        /* CDcf: only the four sphere-slave detectors are on this path */
        if (msl->shape != VO_SPHERE || (mms->shape != VO_SPHERE && mms->shape != VO_QUASIPLATE &&
                                        mms->shape != VO_QUASICYLINDER && mms->shape != VO_QUASIPLATE_HOLE))
            continue;
        iact *it = &s->ia[s->n_ia++];
        it->slave = sl; it->master = ms; it->lo = p1; it->hi = p2;
        it->law = im->law; it->p = im->p;
        const record *r = find_rec(s, p1, p2);
        if (r) it->st = r->st; else status_default(&it->st);
        detect(s, it);
    }
    return 0;
}

/* RebuildIactRecordTab (LeapConsultant.cpp:264-279 / NBSConsultant.cpp:376-395) */
static int rebuild_records(vo_sim *s)
{
    int n = 0;
    for (int k = 0; k < s->n_ia; ++k) n += s->ia[k].ci.active;
    free(s->rec);
    s->rec = malloc(sizeof(record) * (n ? n : 1));
    if (!s->rec) return -1;
    s->n_rec = 0;
    for (int k = 0; k < s->n_ia; ++k)
        if (s->ia[k].ci.active) {
            record *r = &s->rec[s->n_rec++];
            r->lo = s->ia[k].lo; r->hi = s->ia[k].hi; r->st = s->ia[k].st;
            r->order = s->n_rec;
        }
    /* IactRecordTab::PushRecord is std::map::insert (IactRecordTab.cpp:36-39):
       a key listed twice (the periodic traversal can reach a cell pair twice)
       keeps its first insertion */
    qsort(s->rec, s->n_rec, sizeof(record), cmp_rec_first);
    int u = 0;
    for (int k = 0; k < s->n_rec; ++k)
        if (u == 0 || s->rec[k].lo != s->rec[u - 1].lo || s->rec[k].hi != s->rec[u - 1].hi) s->rec[u++] = s->rec[k];
    s->n_rec = u;
    return 0;
}

/* Boundary::InBoundary (Boundary.cpp:161-191) */
static int in_zoi(const vo_sim *s, v3 p)
{
    const double c[3] = {p.x, p.y, p.z};
    for (int d = 0; d < 3; ++d)
        if (s->zoi_on[d] && (c[d] < s->zoi_lo[d] || c[d] > s->zoi_hi[d])) return 0;
    return 1;
}

static int cmp_rec_pos(const void *pa, const void *pb)
{
    /* by key, then by position in the source table (the later assignment wins) */
    const record *a = pa, *b = pb;
    int c = cmp_rec(pa, pb);
    if (c) return c;
    return (a->order > b->order) - (a->order < b->order);
}

/* Consultant::CleanUp (Consultant.cpp:264-319): mobile DOs outside the zone
 * of interest are erased; the remaining DO ids close up, and the
 * IactRecordTab keys move with IactRecordTab::ModifyPair
 * (IactRecordTab.cpp:41-69).  ModifyPair moves every key by the mapping
 * id -> id + jump[id], jump[id] being minus the number of erased DOs up to
 * and including id (unsigned arithmetic): records of erased DOs are not
 * dropped but move to their predecessor's new id; moved keys are
 * collected in a std::map (the later of two equal keys wins, in key order)
 * and written over the table's entries with the same keys. */
static int cleanup(vo_sim *s)
{
    if (!s->zoi_active) return 0;
    const int n = s->n_do;
    unsigned *jump = malloc(sizeof(unsigned) * (n ? n : 1));
    char *dead = calloc(n ? n : 1, 1);
    if (!jump || !dead) { free(jump); free(dead); return -1; }
    unsigned ul_dead = 0;
    int nd = 0;
    for (int i = 0; i < n; ++i) {
        if (s->models[s->model_of[i]].behavior == VO_MOBILE && !in_zoi(s, s->pos[i])) {
            dead[i] = 1;
            ++nd;
            --ul_dead;
        }
        jump[i] = ul_dead;
    }
    if (nd == 0) { free(jump); free(dead); return 0; }
    /* ModifyPair */
    record *moved = malloc(sizeof(record) * (s->n_rec ? s->n_rec : 1));
    record *kept = malloc(sizeof(record) * (s->n_rec ? s->n_rec : 1));
    if (!moved || !kept) { free(jump); free(dead); free(moved); free(kept); return -1; }
    int nm = 0, nk = 0;
    for (int k = 0; k < s->n_rec; ++k) {   /* the table is in key order */
        const record *r = &s->rec[k];
        const unsigned lo = r->lo + jump[r->lo], hi = r->hi + jump[r->hi];
        if (lo != r->lo || hi != r->hi) {
            s->cu_moved++;
            s->cu_from_dead += dead[r->lo] || dead[r->hi];
            moved[nm] = *r;
            moved[nm].lo = lo; moved[nm].hi = hi;
            moved[nm].order = nm;
            ++nm;
        } else {
            kept[nk++] = *r;
        }
    }
    qsort(moved, nm, sizeof(record), cmp_rec_pos);
    int nu = 0;   /* the last of each run of equal keys */
    for (int k = 0; k < nm; ++k) {
        if (k + 1 < nm && moved[k + 1].lo == moved[k].lo && moved[k + 1].hi == moved[k].hi) continue;
        moved[nu++] = moved[k];
    }
    /* merge: a moved record replaces a kept one with the same key */
    int nr = 0, a = 0, b = 0;
    while (a < nk || b < nu) {
        if (b >= nu || (a < nk && cmp_rec(&kept[a], &moved[b]) < 0)) s->rec[nr++] = kept[a++];
        else {
            if (a < nk && cmp_rec(&kept[a], &moved[b]) == 0) {
                ++a;
                s->cu_replaced++;
            }
            s->rec[nr++] = moved[b++];
        }
    }
    s->n_rec = nr;
    free(moved); free(kept);
    /* DOContainer::Erase + DOWorld::EraseDOStatus: the DOs close up */
    int w = 0;
    for (int i = 0; i < n; ++i) {
        if (dead[i]) continue;
        s->pos[w] = s->pos[i]; s->vel[w] = s->vel[i]; s->ox[w] = s->ox[i]; s->oz[w] = s->oz[i];
        s->av[w] = s->av[i]; s->imp[w] = s->imp[i]; s->aimp[w] = s->aimp[i];
        s->last_imp[w] = s->last_imp[i]; s->last_aimp[w] = s->last_aimp[i];
        s->do_sudo[w] = s->do_sudo[i]; s->model_of[w] = s->model_of[i]; s->orig_id[w] = s->orig_id[i];
        ++w;
    }
    s->cu_erased += n - w;
    s->n_do = w;
    /* the interactions and the candidate list are rebuilt by the Reset that follows */
    s->n_ia = 0;
    s->n_cand = 0;
    free(jump); free(dead);
    return 0;
}

static int consultant_reset(vo_sim *s)
{
    return s->consultant == 1 ? reset_nbs(s) : reset_leap(s);
}

vo_sim *vo_create(const vo_scene *sc, char *err, int errlen)
{
    vo_sim *s = calloc(1, sizeof(vo_sim));
    if (!s) return NULL;
    int n = sc->n_do;
    s->n_do = n; s->n_models = sc->n_models; s->n_groups = sc->n_groups;
    s->models = malloc(sizeof(vo_model) * sc->n_models);
    s->iact = malloc(sizeof(vo_iact) * sc->n_groups * sc->n_groups);
    s->model_of = malloc(sizeof(int) * (n ? n : 1));
    s->range = malloc(sizeof(double) * sc->n_models);
    s->mass = malloc(sizeof(double) * sc->n_models);
    s->sudo = malloc(sizeof(double) * sc->n_models);
    v3 **vs[] = {&s->pos, &s->vel, &s->ox, &s->oz, &s->av, &s->imp, &s->aimp, &s->last_imp, &s->last_aimp};
    for (unsigned k = 0; k < sizeof(vs) / sizeof(vs[0]); ++k) *vs[k] = calloc(n ? n : 1, sizeof(v3));
    s->do_sudo = malloc(sizeof(double) * (n ? n : 1));
    s->orig_id = malloc(sizeof(unsigned) * (n ? n : 1));
    if (!s->models || !s->iact || !s->model_of || !s->range || !s->mass || !s->sudo || !s->do_sudo || !s->orig_id ||
        !s->pos || !s->vel || !s->ox || !s->oz || !s->av || !s->imp || !s->aimp || !s->last_imp ||
        !s->last_aimp) {
        snprintf(err, errlen, "out of memory");
        vo_destroy(s);
        return NULL;
    }
    memcpy(s->models, sc->models, sizeof(vo_model) * sc->n_models);
    memcpy(s->iact, sc->iact, sizeof(vo_iact) * sc->n_groups * sc->n_groups);
    memcpy(s->model_of, sc->model_of, sizeof(int) * n);
    for (int m = 0; m < sc->n_models; ++m) {
        s->range[m] = model_range(&s->models[m]);
        model_mass(&s->models[m], &s->mass[m], &s->sudo[m]);
    }
    for (int i = 0; i < n; ++i) {
        s->pos[i] = ld(sc->pos + 3 * i); s->vel[i] = ld(sc->vel + 3 * i);
        s->ox[i] = ld(sc->ox + 3 * i); s->oz[i] = ld(sc->oz + 3 * i); s->av[i] = ld(sc->av + 3 * i);
        const vo_model *m = &s->models[s->model_of[i]];
        /* fixed/orbital wall ctors set dSudoMass = 0 (DOFixedQuasiPlate.cpp:12) */
        if (m->behavior == VO_FIXED && m->shape != VO_SPHERE) s->do_sudo[i] = 0.0;
        else if (m->behavior == VO_ORBITAL) s->do_sudo[i] = 0.0;
        else s->do_sudo[i] = s->sudo[s->model_of[i]];
    }
    s->dt = sc->dt; s->t_start = sc->t_start; s->t_stop = sc->t_stop; s->t_current = sc->t_current;
    s->g = ld(sc->g); s->sf = sc->safety_factor;
    s->up_iact = sc->up_iact ? sc->up_iact : 1; s->record_every = sc->record ? sc->record : 0xffffffffu;
    s->consultant = sc->consultant;
    for (int i = 0; i < n; ++i) s->orig_id[i] = (unsigned)i;
    /* Boundary::Correct: lower/upper sorted per axis, an axis bounded where they differ */
    for (int d = 0; d < 3; ++d) {
        s->zoi_lo[d] = fmin(sc->zoi_lo[d], sc->zoi_hi[d]);
        s->zoi_hi[d] = fmax(sc->zoi_lo[d], sc->zoi_hi[d]);
        s->zoi_on[d] = s->zoi_lo[d] != s->zoi_hi[d];
    }
    s->zoi_active = s->zoi_on[0] || s->zoi_on[1] || s->zoi_on[2];
    for (int d = 0; d < 3; ++d) {
        s->pbc.lo[d] = fmin(sc->pbc_lo[d], sc->pbc_hi[d]);
        s->pbc.hi[d] = fmax(sc->pbc_lo[d], sc->pbc_hi[d]);
        s->pbc.len[d] = s->pbc.hi[d] - s->pbc.lo[d];
        s->pbc.on[d] = s->pbc.lo[d] != s->pbc.hi[d];
    }
    s->pbc.active = s->pbc.on[0] || s->pbc.on[1] || s->pbc.on[2];
    if (s->zoi_active && s->consultant != 0) {
        snprintf(err, errlen, "zone of interest: only the LeapConsultant path is restated");
        vo_destroy(s);
        return NULL;
    }
    if (consultant_reset(s) || initiate(s)) {
        snprintf(err, errlen, "out of memory in consultant reset");
        vo_destroy(s);
        return NULL;
    }
    /* SimMediator::SimMediator (SimMediator.cpp:122-134): CleanUp after the
       first Initiate; a scene that lost DOs is reset and initiated again */
    const int before = s->n_do;
    if (cleanup(s) || (s->n_do != before && (consultant_reset(s) || initiate(s)))) {
        snprintf(err, errlen, "out of memory in the zone-of-interest clean-up");
        vo_destroy(s);
        return NULL;
    }
    return s;
}

void vo_destroy(vo_sim *s)
{
    if (!s) return;
    free(s->models); free(s->iact); free(s->model_of); free(s->range); free(s->mass); free(s->sudo);
    free(s->pos); free(s->vel); free(s->ox); free(s->oz); free(s->av); free(s->imp); free(s->aimp);
    free(s->last_imp); free(s->last_aimp); free(s->do_sudo); free(s->orig_id);
    free(s->cand_lo); free(s->cand_hi); free(s->ia); free(s->rec);
    free(s);
}

int vo_set_history(vo_sim *s, int n, const unsigned *lo, const unsigned *hi, const vo_status *st)
{
    free(s->rec);
    s->rec = malloc(sizeof(record) * (n ? n : 1));
    if (!s->rec) return -1;
    for (int k = 0; k < n; ++k) { s->rec[k].lo = lo[k]; s->rec[k].hi = hi[k]; s->rec[k].st = st[k]; }
    s->n_rec = n;
    qsort(s->rec, n, sizeof(record), cmp_rec);
    return initiate(s);
}

/* DOSphere::Response (DOSphere.cpp:21-48) and the wall responses
 * (DOFixedQuasiPlate.h:16-19, DOOrbitalQuasiPlate.cpp:18-34,
 * DOConstrainedQuasiPlate.cpp:18-35, DOOrbitalQuasiCylinder.cpp:17-33). */
static void response(vo_sim *s, int i)
{
    const vo_model *m = &s->models[s->model_of[i]];
    double dt = s->dt;
    if (m->shape == VO_SPHERE && m->behavior == VO_MOBILE) {
        v3 V = s->vel[i], AV = s->av[i], P = s->pos[i], Ox = s->ox[i], Oz = s->oz[i];
        double r = m->attr[0], sm = s->do_sudo[i];
        v3 dv = smul(1.0 / sm, s->imp[i]);
        v3 dav = smul(2.5 / sm / r / r, s->aimp[i]);
        v3 dp = smul(dt, add(V, smul(0.5, dv)));
        v3 dw = smul(dt, add(AV, smul(0.5, dav)));
        s->pos[i] = add(P, dp);
        s->ox[i] = rotate_around(Ox, dw);
        s->oz[i] = rotate_around(Oz, dw);
        s->vel[i] = add(V, dv);
        s->av[i] = add(AV, dav);
    } else if (m->behavior == VO_ORBITAL && m->shape != VO_SPHERE) {
        v3 dp = muls(s->vel[i], dt), dw = muls(s->av[i], dt);
        s->pos[i] = add(s->pos[i], dp);
        s->ox[i] = rotate_around(s->ox[i], dw);
        s->oz[i] = rotate_around(s->oz[i], dw);
    } else if (m->behavior == VO_CONSTRAINED && m->shape != VO_SPHERE) {
        v3 V = s->vel[i], Oz = s->oz[i];
        v3 dv = smul(1.0 / s->do_sudo[i], smul(dot(s->imp[i], Oz), dir(Oz)));
        v3 dp = smul(dt, add(V, smul(0.5, dv)));
        v3 dw = muls(s->av[i], dt);
        s->pos[i] = add(s->pos[i], dp);
        s->ox[i] = rotate_around(s->ox[i], dw);
        s->oz[i] = rotate_around(Oz, dw);
        s->vel[i] = add(V, dv);
    }
    /* fixed: ClearImpact only */
    s->last_imp[i] = s->imp[i]; s->last_aimp[i] = s->aimp[i];
    s->imp[i] = mk(0, 0, 0); s->aimp[i] = mk(0, 0, 0);
}

static int step_impl(vo_sim *s, int nas, const unsigned *ids, const double *apos, const double *avel,
                     const double *aox, const double *aoz, const double *aav)
{
    double dt = s->dt;
    /* 1. IactContainer::CalculateImpact (IactContainer.cpp:33-37) */
    for (int k = 0; k < s->n_ia; ++k) {
        iact *it = &s->ia[k];
        unsigned a = it->slave, b = it->master;
        body A = {s->pos[a], s->vel[a], s->av[a], s->do_sudo[a]};
        body B = {s->pos[b], s->vel[b], s->av[b], s->do_sudo[b]};
        solve_out o;
        solve(it->law, it->p, &A, &B, &it->ci, dt, sim_bc(s), &it->st, &o);
        if (o.added) {
            /* DiscreteObject::AddImpact (DiscreteObject.cpp:44-53) */
            s->imp[a] = add(s->imp[a], o.imp_a);
            s->imp[b] = add(s->imp[b], o.linear_only ? neg(o.imp_a) : smul(-1.0, o.imp_a));
            if (!o.linear_only) {
                s->aimp[a] = add(s->aimp[a], o.ang_a);
                s->aimp[b] = add(s->aimp[b], o.ang_b);
            }
        }
    }
    /* 3. AddFieldImpact + AddConstrainedImpact (DOContainer.cpp:20-50) */
    v3 gdt = muls(s->g, dt);
    for (int i = 0; i < s->n_do; ++i) s->imp[i] = add(s->imp[i], muls(gdt, s->do_sudo[i]));
    for (int i = 0; i < s->n_do; ++i) {
        const vo_model *m = &s->models[s->model_of[i]];
        s->imp[i] = add(s->imp[i], smul(dt, ld(m->fext)));
    }
    /* 4. Response */
    for (int i = 0; i < s->n_do; ++i) response(s, i);
    for (int k = 0; k < nas; ++k) {
        unsigned i = ids[k];
        if ((int)i >= s->n_do) return -1;
        s->pos[i] = ld(apos + 3 * k); s->vel[i] = ld(avel + 3 * k);
        s->ox[i] = ld(aox + 3 * k); s->oz[i] = ld(aoz + 3 * k); s->av[i] = ld(aav + 3 * k);
        /* DOContainer::SetDOStatus replaces the whole DOStatus, including the
         * stored last impulses, which an assigned status carries as zero */
        s->last_imp[i] = mk(0, 0, 0); s->last_aimp[i] = mk(0, 0, 0);
    }
    /* DOContainer::EnforcePeriodicBoundaryConditions (SimMediator.cpp:746, 889;
       DiscreteObject.cpp:60-65): every DO's position */
    if (s->pbc.active)
        for (int i = 0; i < s->n_do; ++i) s->pos[i] = bc_enforce(&s->pbc, s->pos[i]);
    /* 5. CheckContactStatus */
    for (int k = 0; k < s->n_ia; ++k) detect(s, &s->ia[k]);
    /* 6. Consultant::NextStep (LeapConsultant.cpp:49-97 / NBSConsultant.cpp:35-78) */
    s->round++;
    s->t_current += dt;   /* SystemParameter::NextStep */
    int rebuild = 0, reset = (s->round % s->up_iact) == 0;
    if (reset) {
        if (rebuild_records(s)) return -1;
        rebuild = 1;
        if (cleanup(s)) return -1;   /* LeapConsultant::NextStep (LeapConsultant.cpp:58-68) */
        if (consultant_reset(s)) return -1;
    }
    if ((s->round % s->record_every) == 0 && !rebuild) {
        if (rebuild_records(s)) return -1;
        rebuild = 1;
    }
    /* IactContainer::BackupImpactStatus (IactContainer.cpp:128-131) */
    for (int k = 0; k < s->n_ia; ++k) {
        s->ia[k].st.last_contact = s->ia[k].st.contact;
        s->ia[k].st.last_bond = s->ia[k].st.bond;
    }
    if (s->t_current >= s->t_stop && !rebuild) {
        if (rebuild_records(s)) return -1;
    }
    /* 8. Initiate on reset */
    if (reset && initiate(s)) return -1;
    return s->t_current < s->t_stop;
}

int vo_step(vo_sim *s) { return step_impl(s, 0, NULL, NULL, NULL, NULL, NULL, NULL); }

int vo_step_assigned(vo_sim *s, int n, const unsigned *ids, const double *pos, const double *vel,
                     const double *ox, const double *oz, const double *av)
{
    return step_impl(s, n, ids, pos, vel, ox, oz, av);
}

void vo_get_state(const vo_sim *s, double *pos, double *vel, double *ox, double *oz, double *av,
                  double *impact, double *ang_impact)
{
    for (int i = 0; i < s->n_do; ++i) {
        if (pos) st3(pos + 3 * i, s->pos[i]);
        if (vel) st3(vel + 3 * i, s->vel[i]);
        if (ox) st3(ox + 3 * i, s->ox[i]);
        if (oz) st3(oz + 3 * i, s->oz[i]);
        if (av) st3(av + 3 * i, s->av[i]);
        if (impact) st3(impact + 3 * i, s->last_imp[i]);
        if (ang_impact) st3(ang_impact + 3 * i, s->last_aimp[i]);
    }
}

double vo_time(const vo_sim *s) { return s->t_current; }

int vo_num_do(const vo_sim *s) { return s->n_do; }

void vo_cleanup_stats(const vo_sim *s, long long *out)
{
    out[0] = s->cu_erased; out[1] = s->cu_moved; out[2] = s->cu_from_dead; out[3] = s->cu_replaced;
}

void vo_get_ids(const vo_sim *s, unsigned *ids)
{
    for (int i = 0; i < s->n_do; ++i) ids[i] = s->orig_id[i];
}

/* DOModel::Analysis vMassMomentInertia (DOModel.cpp:22-29, 56-89) */
static v3 model_mmi(const vo_sim *s, int m)
{
    const vo_model *M = &s->models[m];
    const double mass = s->mass[m];
    double r, h, w, l, d;
    switch (M->shape) {
    case VO_SPHERE:
        r = M->attr[0];
        d = 0.4 * mass * r * r;
        return mk(d, d, d);
    case VO_QUASICYLINDER:
        r = M->attr[0]; h = M->attr[1];
        d = mass * (3.0 * r * r + h * h) / 12.0;
        return mk(d, d, 0.5 * mass * r * r);
    case VO_QUASIPLATE:
        w = M->attr[0]; h = M->attr[1]; l = M->attr[2];
        return smul(mass / 12.0, mk(h * h + l * l, w * w + l * l, h * h + w * w));
    default: return mk(0, 0, 0);
    }
}

/* DOWorld::CalculateSystemEnergy (DOWorld.cpp:481-632), statement by
 * statement: the first mobile DO initialises the sums, the others add to
 * them; the averages are 1.0/N times the sums (NaN when there is no mobile
 * DO, as in the reference); the norms are a second sequential pass. */
void vo_system_energy(const vo_sim *s, double *out)
{
    double pe = 0, et = 0, er = 0, vmax = 0, vmin = 0, wmax = 0, wmin = 0, mn = 0, ln = 0;
    v3 P = mk(0, 0, 0), L = mk(0, 0, 0);
    unsigned nm = 0;
    int first = -1;
    for (int i = 0; i < s->n_do; ++i) {
        const int m = s->model_of[i];
        if (s->models[m].behavior != VO_MOBILE) continue;
        const double mass = s->mass[m];
        const v3 mmi = model_mmi(s, m);
        const double v = len(s->vel[i]);
        const v3 av = s->av[i];
        if (first < 0) {
            first = i;
            P = smul(mass, s->vel[i]);
            L = cross(mmi, av);
            pe = -mass * dot(s->g, s->pos[i]);
            et = 0.5 * mass * v * v;
            er = 0.5 * (mmi.x * av.x * av.x + mmi.y * av.y * av.y + mmi.z * av.z * av.z);
            vmax = vmin = v;
            wmax = wmin = len(av);
            nm = 1;
            continue;
        }
        const double dav = len(av);
        nm++;
        P = add(P, smul(mass, s->vel[i]));
        L = add(L, cross(mmi, av));
        pe -= mass * dot(s->g, s->pos[i]);
        et += 0.5 * mass * v * v;
        er += 0.5 * (mmi.x * av.x * av.x + mmi.y * av.y * av.y + mmi.z * av.z * av.z);
        if (v > vmax) vmax = v;
        else if (v < vmin) vmin = v;
        if (dav > wmax) wmax = dav;
        else if (dav < wmin) wmin = dav;
    }
    P = smul(1.0 / (double)nm, P);
    L = smul(1.0 / (double)nm, L);
    if (first >= 0)
        for (int i = first; i < s->n_do; ++i) {
            const int m = s->model_of[i];
            if (s->models[m].behavior != VO_MOBILE) continue;
            mn += len(sub(smul(s->mass[m], s->vel[i]), P));
            ln += len(sub(cross(model_mmi(s, m), s->av[i]), L));
        }
    const double o[16] = {pe, et, er, P.x, P.y, P.z, L.x, L.y, L.z, mn, ln, vmax, vmin, wmax, wmin, (double)nm};
    for (int k = 0; k < 16; ++k) out[k] = o[k];
}

/* ---- Knight analyses of a DOWorld state (src/vedo/knight) ---- */

/* The spheres whose centre lies within the box widened by their radius, all
 * comparisons strict (CalcCoordinateNum.cpp:24-49, CalcDensity.cpp:20-43). */
static int knight_in_box(const vo_sim *s, int i, const double *b)
{
    const vo_model *m = &s->models[s->model_of[i]];
    if (m->shape != VO_SPHERE) return 0;
    const double r = m->attr[0];
    const v3 p = s->pos[i];
    return p.x < b[1] + r && p.x > b[0] - r && p.y < b[3] + r && p.y > b[2] - r && p.z < b[5] + r &&
           p.z > b[4] - r;
}

/* CalcCoordinateNum::computeAvgCoordinateNum (CalcCoordinateNum.cpp:8-77):
 * for every sphere of the box, the other spheres of the whole world whose
 * centre distance |x_i - x_j| is at most r_i + r_j; total over the box's
 * spheres divided by their number (NaN for none). */
double vo_coordination_number(const vo_sim *s, const double *box)
{
    unsigned total = 0, n = 0;
    for (int i = 0; i < s->n_do; ++i) {
        if (!knight_in_box(s, i, box)) continue;
        ++n;
        const double ri = s->models[s->model_of[i]].attr[0];
        unsigned count = 0;
        for (int j = 0; j < s->n_do; ++j) {
            if (j == i) continue;
            const vo_model *mj = &s->models[s->model_of[j]];
            if (mj->shape != VO_SPHERE) continue;
            if (len(sub(s->pos[i], s->pos[j])) <= ri + mj->attr[0]) count++;
        }
        total += count;
    }
    return total / (double)n;
}

/* njr::RandomGenerator (RandomGenerator.cpp:70-126): a subtractive
 * lagged-Fibonacci generator modulo 1e9 over 55 words; sprand(seed), then
 * 165 draws discarded.  GeneratingDouble(lo, hi) = lprand()/1e9*(hi-lo)+lo. */
typedef struct { int a, b, arr[55]; } njr_rng;
static int njr_next(njr_rng *g)
{
    if ((g->a--) == 0) g->a = 54;
    if ((g->b--) == 0) g->b = 54;
    int t = g->arr[g->a] - g->arr[g->b];
    if (t < 0) t += 1000000000;
    g->arr[g->a] = t;
    return t;
}
static void njr_seed(njr_rng *g, int seed)
{
    int last = seed, next = 1;
    g->arr[0] = seed;
    for (int i = 1; i < 55; i++) {
        const int ii = (21 * i) % 55;
        g->arr[ii] = next;
        next = last - next;
        if (next < 0) next += 1000000000;
        last = g->arr[ii];
    }
    g->a = 0;
    g->b = 24;
    for (int i = 0; i < 165; i++) njr_next(g);
}
static double njr_double(njr_rng *g, double lo, double hi)
{
    if (lo > hi) { const double t = lo; lo = hi; hi = t; }
    else if (lo == hi) return hi;
    return (double)njr_next(g) / 1000000000 * (hi - lo) + lo;
}

/* CalcDensity::computeDensity (CalcDensity.cpp:12-72): n_points random
 * points of the box (x, y, z drawn in that order from a generator seeded
 * with `seed`; the reference's first generator of a process uses 999); the
 * percentage of points inside a sphere of the box, formed in float as the
 * reference does (hit / (float)n * 100.0). */
double vo_packing_density(const vo_sim *s, const double *box, int n_points, int seed)
{
    njr_rng g;
    njr_seed(&g, seed);
    unsigned hit = 0;
    for (int k = 0; k < n_points; ++k) {
        const double x = njr_double(&g, box[0], box[1]);
        const double y = njr_double(&g, box[2], box[3]);
        const double z = njr_double(&g, box[4], box[5]);
        const v3 p = mk(x, y, z);
        for (int j = 0; j < s->n_do; ++j) {
            if (!knight_in_box(s, j, box)) continue;
            if (len(sub(p, s->pos[j])) <= s->models[s->model_of[j]].attr[0]) {
                hit++;
                break;
            }
        }
    }
    const float f = (float)hit / (float)n_points;
    return (double)f * 100.0;
}

/* CalculateGranularTemperature (Knight.cpp:1162-1225) over a GSRectangle
 * (GSRectangle.cpp:15-30: centre c, axes Ox, Oz, width/length/height along
 * Ox, Oz x Ox, Oz): the DOs of model `model` inside it, in DO-id order; their
 * mean velocity and angular velocity, and the mean squared deviations / 3.
 * out[5]: count, |mean v|, T_v, |mean w|, T_w. */
void vo_granular_temperature(const vo_sim *s, int model, const double *c, const double *ox, const double *oz,
                             const double *size, double *out)
{
    const v3 C = ld(c), Ox = ld(ox), Oz = ld(oz), Oy = cross(Oz, Ox);
    v3 av = mk(0, 0, 0), aw = mk(0, 0, 0);
    unsigned n = 0;
    char *in = calloc(s->n_do ? s->n_do : 1, 1);
    for (int i = 0; i < s->n_do; ++i) {
        if (s->model_of[i] != model) continue;
        const v3 lp = sub(s->pos[i], C);
        const double lx = dot(lp, Ox), ly = dot(lp, Oy), lz = dot(lp, Oz);
        if (fabs(lx) <= 0.5 * size[0] && fabs(ly) <= 0.5 * size[1] && fabs(lz) <= 0.5 * size[2]) {
            av = add(av, s->vel[i]);
            aw = add(aw, s->av[i]);
            in[i] = 1;
            n++;
        }
    }
    if (n != 0) {
        av = smul(1.0 / (double)n, av);
        aw = smul(1.0 / (double)n, aw);
    }
    double tv = 0.0, tw = 0.0;
    for (int i = 0; i < s->n_do; ++i) {
        if (!in[i]) continue;
        const v3 dv = sub(s->vel[i], av), dw = sub(s->av[i], aw);
        tv += dot(dv, dv);
        tw += dot(dw, dw);
    }
    if (n != 0) {
        tv = tv / (double)n / 3.0;
        tw = tw / (double)n / 3.0;
    }
    free(in);
    out[0] = (double)n; out[1] = len(av); out[2] = tv; out[3] = len(aw); out[4] = tw;
}

int vo_num_interactions(const vo_sim *s) { return s->n_ia; }
void vo_get_interactions(const vo_sim *s, unsigned *slave, unsigned *master)
{
    for (int k = 0; k < s->n_ia; ++k) { slave[k] = s->ia[k].slave; master[k] = s->ia[k].master; }
}
int vo_num_candidates(const vo_sim *s) { return s->n_cand; }
void vo_get_candidates(const vo_sim *s, unsigned *lo, unsigned *hi)
{
    for (int k = 0; k < s->n_cand; ++k) { lo[k] = s->cand_lo[k]; hi[k] = s->cand_hi[k]; }
}

/* The IactRecordTab of the last RebuildIactRecordTab (rebuild or record
 * step; NBSConsultant.cpp:376-395): its keys in key order. */
int vo_num_records(const vo_sim *s) { return s->n_rec; }
void vo_get_records(const vo_sim *s, unsigned *lo, unsigned *hi)
{
    for (int k = 0; k < s->n_rec; ++k) {
        lo[k] = s->rec[k].lo;
        hi[k] = s->rec[k].hi;
    }
}

int vo_num_active(const vo_sim *s)
{
    int n = 0;
    for (int k = 0; k < s->n_ia; ++k) n += s->ia[k].ci.active;
    return n;
}
void vo_get_active(const vo_sim *s, unsigned *lo, unsigned *hi, vo_status *st)
{
    int n = 0;
    for (int k = 0; k < s->n_ia; ++k)
        if (s->ia[k].ci.active) {
            lo[n] = s->ia[k].lo; hi[n] = s->ia[k].hi;
            if (st) st[n] = s->ia[k].st;
            n++;
        }
}
