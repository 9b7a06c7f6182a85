// vg_device.h -- device data layout and per-contact physics of the DEM step.
//
// Layout in HBM (the spheres of every scene of a batch, scene by scene, each
// scene's in cell order, re-sorted at every candidate-list rebuild; arrays
// have a fixed leading dimension LD >= N, the batch's sphere capacity):
//   pos  double4 [LD]  x, y, z, radius
//   vel  double4 [LD]  vx, vy, vz, path length since the last neighbour relist
//   avl  double4 [LD]  wx, wy, wz, sudo-mass m* (DOModel.cpp:22-30)
//   ori  double  [6][LD]  Ox.xyz, Oz.xyz (orientation, SoA)
//   oid  u32 [LD]  sphere rank within its scene (order of DO ids)
//   mdl  i32 [LD]  DOModel index (batch-wide table; the model names the scene)
// Neighbour ("inner") list, ELL transposed so that slot k of 64 consecutive
// spheres is one contiguous 256-byte run:
//   hdr  u8  [LD]          count (7 bits) | current half (bit 7)
//   ent  u32 [2][KI][LD]   partner (26 bits: sphere slot, or LD + wall slot) | status flags
//   slot u8  [2][KI][LD]   history slot of an entry whose Contact flag is set
//   hd   double4 [KI][LD]  per history slot: tangential force history xyz,
//                          Kn (ISwH) / initial velocity (ISwHertz)
// A contact keeps its history slot while it lasts; a new contact takes the
// sphere's lowest free slot, so a sphere's histories fill the first rows and
// the histories of neighbouring spheres share cache lines.
// Every sphere keeps its own copy of each pair's ImpactStatus; the two copies
// are updated by the same instruction sequence from the same inputs, so they
// stay bit-identical and no thread ever reads another thread's history.
#pragma once

#include "vg_math.h"

namespace vg {

constexpr unsigned PARTNER_MASK = (1u << 26) - 1;
// Candidate-list entries carry two marks of the rebuild that listed them:
// the pair is within the scene's skin of contact (depth >= -skin) and in
// contact (depth > 0) at the rebuild's positions.
constexpr unsigned NBR_NEAR = 1u << 31;
constexpr unsigned NBR_ACTIVE = 1u << 30;
constexpr unsigned F_CONTACT = 1u << 26;
constexpr unsigned F_LASTCONTACT = 1u << 27;
constexpr unsigned F_BOND = 1u << 28;
constexpr unsigned F_LASTBOND = 1u << 29;
constexpr unsigned F_SLAVE = 1u << 30;   // this sphere is the pair's slave (lower DO id)
constexpr unsigned F_NOBACKUP = 1u << 31;   // status restored at a rebuild: skip the pending Backup
constexpr unsigned F_STATUS = F_CONTACT | F_LASTCONTACT | F_BOND | F_LASTBOND;
// ImpactStatus() defaults: bLastContact(true), bLastBond(true) (ImpactStatus.cpp:6-9)
constexpr unsigned F_DEFAULT = F_LASTCONTACT | F_LASTBOND | F_NOBACKUP;

enum { LAW_NONE = 0, BSDBF = 1, BSDB2F = 2, BTSDBF = 3, LAW_H = 4, LAW_HERTZ = 5 };
enum { SH_SPHERE = 1, SH_CYL = 6, SH_PLATE = 7 };

struct LawDev {
    int law;
    int pad;
    double p[8];
};

struct ModelDev {
    int group;   // DOGroup index within the model's scene
    int scene;   // scene slot of the batch
    double fext[3];
};

struct WallDev {
    d3 pos, vel, ox, oz, av;
    double a0, a1, a2;   // plate: w, h, l ; cylinder: r, h
    double range;        // DOModel::Analysis range (global scope)
    double reach;        // distance from the centre to the farthest point of the wall
    double path;         // bound on the displacement of the wall's contact rectangle / axis segment since the last relist
    double msudo;        // sudo-mass: DensityFactor * mass for a constrained wall, 0 for fixed / orbital
    double fext[3];      // DOModel external force (AddConstrainedImpact)
    int shape, group;
    int behavior, kin;   // kin: 0 none, 1 host-assigned, 2 device rotation
    int scene;           // scene slot of the batch
    int constrained;     // DOConstrained* response (moves along Oz under its impulse)
};

struct Contact {
    d3 n, p;
    double depth;
    bool active;
};

struct Hist {
    unsigned f;   // status flags
    d3 s;         // shear force
    double aux;   // Kn (ISwH) or initial velocity (ISwHertz)
};

// SystemParameter PeriodicBoundaryConditions after Boundary::Correct
// (Boundary.cpp:79-129): an axis is periodic where lower != upper, len =
// upper - lower.  pax / maxid: the axes the Leap traversal wraps at the last
// rebuild (periodic and spanned by the spheres for at least half their range,
// LeapConsultant.h:146-150) and the last zone index per axis.
struct PbcDev {
    int on[3];
    int pax[3];
    unsigned maxid[3];
    int pad;
    double lo[3], hi[3], len[3];
};

// Boundary::DifferenceBoundaryConditions (Boundary.cpp:303-349)
__device__ inline double bc_diff1(double d, double l)
{
    if (d >= 0.5 * l) d -= l;
    if (d < -0.5 * l) d += l;
    return d;
}
__device__ inline d3 bc_diff(const PbcDev *b, d3 d)
{
    if (!b) return d;
    if (b->on[0]) d.x = bc_diff1(d.x, b->len[0]);
    if (b->on[1]) d.y = bc_diff1(d.y, b->len[1]);
    if (b->on[2]) d.z = bc_diff1(d.z, b->len[2]);
    return d;
}
// Boundary::EnforceBoundaryConditions (Boundary.cpp:225-301)
__device__ inline double bc_wrap1(double x, double lo, double hi, double l)
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
__device__ inline d3 bc_enforce(const PbcDev *b, d3 x)
{
    if (!b) return x;
    if (b->on[0]) x.x = bc_wrap1(x.x, b->lo[0], b->hi[0], b->len[0]);
    if (b->on[1]) x.y = bc_wrap1(x.y, b->lo[1], b->hi[1], b->len[1]);
    if (b->on[2]) x.z = bc_wrap1(x.z, b->lo[2], b->hi[2], b->len[2]);
    return x;
}

struct PairOut {
    d3 imp, ang_a, ang_b;
    bool added, linear_only;
};

struct RecOut {   // ImpactStatus fields only the record path stores
    d3 point, dir, to_master, ang_to_master;
    double overlap;
};

// CDSphere_Sphere::CalDistance + Detect (CDSphere2Sphere.cpp:14-60)
__device__ inline void cd_ss(d3 xa, double ra, d3 xb, double rb, Contact &c)
{
    d3 im = xb - xa;
    c.depth = ra + rb - len(im);
    c.n = dir(im);
    c.p = xa + smul(ra - 0.5 * c.depth, c.n);
    c.active = c.depth > 0;
}
// the same under a periodic boundary: the centre-to-centre vector as its
// minimum image (CDSphere2Sphere.cpp:19-22)
__device__ inline void cd_ss_bc(d3 xa, double ra, d3 xb, double rb, const PbcDev *b, Contact &c)
{
    d3 im = bc_diff(b, xb - xa);
    c.depth = ra + rb - len(im);
    c.n = dir(im);
    c.p = xa + smul(ra - 0.5 * c.depth, c.n);
    c.active = c.depth > 0;
}

// CDSphere_QuasiPlate::CalDistance + Detect (CDSphere2QuasiPlate.cpp:15-97)
__device__ inline void cd_sp(d3 xa, double ra, const WallDev &w, Contact &c)
{
    d3 oy = cross(w.oz, w.ox);
    d3 cap = xa - project_on(xa - w.pos, w.oz);
    double hh = 0.5 * w.a1, hw = 0.5 * w.a0, hl = 0.5 * w.a2;
    double dx = dot(cap - w.pos, w.ox);
    if (dx < -hw) dx = -hw;
    else if (dx > hw) dx = hw;
    double dy = dot(cap - w.pos, oy);
    if (dy < -hl) dy = -hl;
    else if (dy > hl) dy = hl;
    d3 caps = (muls(w.ox, dx) + muls(oy, dy)) + w.pos;
    d3 im = caps - xa;
    c.depth = ra + hh - len(im);
    c.n = dir(im);
    c.p = xa + smul(ra - 0.5 * c.depth, c.n);
    c.active = c.depth > 0;
}

// CDSphere_QuasiCylinder::CalDistance + Detect (CDSphere2QuasiCylinder.cpp:16-80)
__device__ inline void cd_sc(d3 xa, double ra, const WallDev &w, Contact &c)
{
    double hh = 0.5 * w.a1;
    double d = dot(xa - w.pos, w.oz);
    if (d < -hh) d = -hh;
    else if (d > hh) d = hh;
    d3 caps = w.pos + smul(d, w.oz);
    d3 im = caps - xa;
    c.depth = ra + w.a0 - len(im);
    c.n = dir(im);
    c.p = xa + smul(ra - 0.5 * c.depth, c.n);
    c.active = c.depth > 0.0;
}

__device__ inline void cd_wall(d3 xa, double ra, const WallDev &w, Contact &c)
{
    if (w.shape == SH_PLATE) cd_sp(xa, ra, w, c);
    else cd_sc(xa, ra, w, c);
}

// CDSphere_QuasiPlateWithCircularHole::CalDistance + Detect
// (CDSphere2QuasiPlateWithCircularHole.cpp:16-134).  The plate's rectangle
// as in cd_sp; where the sphere's centre projects into the hole, the closest
// point moves out to the hole's rim.  hole = (radius, x offset, y offset,
// 1): the hole's centre is the plate's centre plus (x, y, 0) in global axes
// (:46), its rim point is placed along the plate's own axes.
__device__ inline void cd_sph(d3 xa, double ra, const WallDev &w, double4 hole, Contact &c)
{
    const double hr = hole.x, hx = hole.y, hy = hole.z;
    d3 oy = cross(w.oz, w.ox);
    d3 cap = xa - project_on(xa - w.pos, w.oz);
    double hh = 0.5 * w.a1, hw = 0.5 * w.a0, hl = 0.5 * w.a2;
    const d3 cha = xa - (w.pos + mk(hx, hy, 0.0));
    double dx = dot(cap - w.pos, w.ox);
    if (dx < -hw) dx = -hw;
    else if (dx > hw) dx = hw;
    double dy = dot(cap - w.pos, oy);
    if (dy < -hl) dy = -hl;
    else if (dy > hl) dy = hl;
    const double dhx = dot(cha, w.ox), dhy = dot(cha, oy);
    const double dh = sqrt(dhx * dhx + dhy * dhy);
    if (dh < hr) {
        if (dh == 0.0) {
            dx = hx;
            dy = hy;
        } else {
            dx = hx + dhx / dh * hr;
            dy = hy + dhy / dh * hr;
        }
    }
    d3 caps = (muls(w.ox, dx) + muls(oy, dy)) + w.pos;
    d3 im = caps - xa;
    if (dh == 0.0) {
        const double l = len(im);
        c.depth = ra + hh - sqrt(l * l + hr * hr);
    } else {
        c.depth = ra + hh - len(im);
    }
    c.n = dir(im);
    c.p = xa + smul(ra - 0.5 * c.depth, c.n);
    c.active = c.depth > 0;
}

// Contact with wall slot ws.  Kernels built for the extended geometry (HOLE)
// look the slot up in the hole table where the handle has one (hole.w = 1
// marks a holed plate, which the wall table stores as a plate); the others
// never read it.
template <bool HOLE>
__device__ inline void cd_wall_at(d3 xa, double ra, const WallDev &w, const double4 *holes, int ws, Contact &c)
{
    if (HOLE && holes) {
        const double4 hole = holes[ws];
        if (hole.w != 0.0) {
            cd_sph(xa, ra, w, hole, c);
            return;
        }
    }
    cd_wall(xa, ra, w, c);
}

// DOMap::CalDistance for sphere-plate / sphere-cylinder (DOMap.cpp:353-412)
__device__ inline double domap_wall_distance(d3 ca, const WallDev &w)
{
    if (w.shape == SH_PLATE) {
        d3 oy = cross(w.oz, w.ox);
        d3 cap = ca - project_on(ca - w.pos, w.oz);
        double hw = 0.5 * w.a0, hl = 0.5 * w.a2;
        double dx = dot(cap - w.pos, w.ox);
        if (dx < -hw) dx = -hw;
        else if (dx > hw) dx = hw;
        double dy = dot(cap - w.pos, oy);
        if (dy < -hl) dy = -hl;
        else if (dy > hl) dy = hl;
        d3 caps = (muls(w.ox, dx) + muls(oy, dy)) + w.pos;
        return len(caps - ca);
    }
    double hh = 0.5 * w.a1;
    double d = dot(ca - w.pos, w.oz);
    d3 cap = w.pos + muls(w.oz, d);
    d3 im;
    if (d < hh && d > -hh) im = cap - ca;
    else if (d >= hh) im = (w.pos + muls(w.oz, hh)) - ca;
    else im = (w.pos - muls(w.oz, hh)) - ca;
    return len(im);
}

// Entry flags as IactContainer::BackupImpactStatus leaves them after the step
// (IactContainer.cpp:128-131, ImpactStatus::Backup): an entry stores the
// Backup of its last solve as pending (applied when it is next solved), so
// readers of a stepped entry apply it; a restored entry (F_NOBACKUP) has no
// pending Backup.
__host__ __device__ inline unsigned backed_up(unsigned f)
{
    if (f & F_NOBACKUP) return f;
    unsigned g = f & ~(F_LASTCONTACT | F_LASTBOND);
    if (f & F_CONTACT) g |= F_LASTCONTACT;
    if (f & F_BOND) g |= F_LASTBOND;
    return g;
}

// ---- ImpactStatus transitions (ImpactStatus.cpp:248-274) ----
__device__ inline void st_clean(Hist &h)
{
    h.f &= ~F_STATUS;
    h.s = mk(0, 0, 0);
    h.aux = 0.0;
}
__device__ inline void st_clean_contact(Hist &h)
{
    unsigned f = h.f & ~F_STATUS;
    if (h.f & F_CONTACT) f |= F_LASTCONTACT;
    if (h.f & F_BOND) f |= F_LASTBOND;
    h.f = f;
    h.s = mk(0, 0, 0);
}

// ---- ISwModels (ISwModels.cpp) ----
__device__ inline d3 normal_force(double kn, double cn, double depth, d3 n, d3 vn)
{
    return smul(-kn * depth, n) - smul(cn, vn);
}
__device__ inline d3 normal_force_bond(double kn, double cn, double depth, d3 n, d3 vn)
{
    if (dot(vn, n) < 0.0) return smul(-kn * depth, n) - smul(cn, vn);
    return smul(-cn, vn);
}
__device__ inline d3 shear_rotation(d3 f, d3 d)
{
    if (dot(f, d) <= 0.0) return smul(-len(f), dir(d));
    return smul(len(f), dir(d));
}

// Law sets a step kernel is specialised for: only the spring-dashpot family,
// only ISwH, or all five (smaller kernels where fewer laws are present);
// LS_EXT is all five laws with the extended geometry: plates with a circular
// hole and periodic boundaries (the only set whose kernels evaluate them).
enum { LS_SD = 0, LS_H = 1, LS_ALL = 2, LS_EXT = 3 };

// One ImpactSolver::NextStep for slave A and master B, for the laws of set LS.
// ISwBSDBF/ISwBSDB2F/ISwBtSDBF (ISwBSDBF.cpp:26-139, ISwBSDB2F.cpp:26-149,
// ISwBtSDBF.cpp:26-137), ISwH (ISwH.cpp:30-156), ISwHertz (ISwHertz.cpp:20-94).
// Under a periodic boundary (LS_EXT with pb) the lever arms are minimum
// images (ISwBSDBF.cpp:69-73, ISwH.cpp:69-73, ISwHertz.cpp:59-63).
template <bool REC, int LS>
__device__ inline void solve(const LawDev &L, double ma, double mb, d3 xa, d3 va, d3 wa, d3 xb, d3 vb,
                             d3 wb, const Contact &c, double dt, Hist &h, PairOut &o, RecOut &r,
                             const PbcDev *pb = nullptr)
{
    o.added = false;
    o.linear_only = false;
    const int law = L.law;
    const double depth = c.depth;
    const d3 n = c.n;
    if (LS >= LS_ALL && law == LAW_HERTZ) {
        o.linear_only = true;
        if (!c.active) {
            st_clean(h);
            return;
        }
        d3 fn = mk(0, 0, 0);
        if (depth > 0.0) {
            d3 ra = c.p - xa, rb = c.p - xb;
            if constexpr (LS == LS_EXT) {
                ra = bc_diff(pb, ra);
                rb = bc_diff(pb, rb);
            }
            d3 rv = ((va - vb) + cross(wa, ra)) - cross(wb, rb);
            d3 vn = smul(dot(rv, n), n);
            if (!(h.f & F_CONTACT)) {
                h.f |= F_CONTACT;
                h.aux = len(vn);
            }
            fn = smul(-L.p[0] * pow(depth, 1.5), n) - smul(L.p[1], vn);
        } else {
            st_clean_contact(h);
        }
        o.imp = muls(fn, dt);
        o.ang_a = o.ang_b = mk(0, 0, 0);
        o.added = true;
        if (REC) {
            r.point = c.p; r.dir = c.n; r.overlap = c.depth;
            r.to_master = -o.imp; r.ang_to_master = mk(0, 0, 0);
        }
        return;
    }
    if (LS != LS_SD && law == LAW_H) {
        const double knh = L.p[0], meff = L.p[1], ksc1 = L.p[2], ksc2 = L.p[3], betan = L.p[4], fk = L.p[5];
        if (!c.active) {
            st_clean(h);
            return;
        }
        d3 ra = mk(0, 0, 0), rb = mk(0, 0, 0), fn = mk(0, 0, 0), fs = mk(0, 0, 0);
        if (depth > 0.0) {
            ra = c.p - xa;
            rb = c.p - xb;
            if constexpr (LS == LS_EXT) {
                ra = bc_diff(pb, ra);
                rb = bc_diff(pb, rb);
            }
            d3 rv = ((va - vb) + cross(wa, ra)) - cross(wb, rb);
            d3 vn = smul(dot(rv, n), n);
            d3 vs = rv - vn;
            double kn;
            if (!(h.f & F_CONTACT)) {
                // p[6]: the law's factor of |vn|^0.4, set on the host
                if (len(vn) == 0.0) kn = knh;
                else kn = L.p[6] * pow(len(vn), 0.4);
                h.aux = kn;
                h.f |= F_CONTACT;
            } else {
                kn = h.aux;
            }
            double ncd = 2.0 * sqrt(meff * kn);
            // ISwH.cpp:117 takes pow(x, 1.0/3.0); cbrt(x) is the same value to a few
            // ulps (the exponent's rounding moves it by a relative 1.9e-17*|ln x|)
            // and costs a fraction of a device pow
            double ks = ksc1 * cbrt(ksc2 * fabs(kn * depth));
            fn = normal_force(kn, ncd * betan, depth, n, vn);
            double ff = fk * kn * depth;
            fs = shear_rotation(h.s, vs) - smul(ks * dt, vs);
            if (len(fs) >= ff) fs = smul(ff, dir(fs));
            h.s = fs;
        } else {
            st_clean_contact(h);
        }
        d3 f = fn + fs;
        o.imp = muls(f, dt);
        o.ang_a = smul(dt, cross(ra, f));
        o.ang_b = smul(-dt, cross(rb, f));
        o.added = true;
        if (REC) {
            r.point = c.p; r.dir = c.n; r.overlap = c.depth;
            r.to_master = smul(-1.0, o.imp); r.ang_to_master = o.ang_b;
        }
        return;
    }
    if constexpr (LS == LS_H) return;
    // spring-dashpot family
    const double kn = L.p[0], ks = L.p[1], bn = L.p[4], bs = L.p[5], fc = L.p[6];
    double me;   // ISwModels::CriticalDamping (ISwModels.cpp:88-108)
    if (mb == 0) me = ma;
    else if (ma == 0) me = mb;
    else me = 0.5 * (ma + mb);
    const double cn = L.p[2] * (2.0 * sqrt(me * kn));
    const double cs = L.p[3] * (2.0 * sqrt(me * ks));
    const bool nbond = (depth <= 0.0) && (kn != 0.0) && (bn != 0.0) && (-depth * kn < bn);
    if (!c.active && !nbond) {
        st_clean(h);
        return;
    }
    d3 ra = mk(0, 0, 0), rb = mk(0, 0, 0), fn = mk(0, 0, 0), fs = mk(0, 0, 0);
    if (depth > 0.0) {
        ra = c.p - xa;
        rb = c.p - xb;
        if constexpr (LS == LS_EXT) {
            ra = bc_diff(pb, ra);
            rb = bc_diff(pb, rb);
        }
        d3 rv = ((va - vb) + cross(wa, ra)) - cross(wb, rb);
        d3 vn = smul(dot(rv, n), n);
        d3 vs = rv - vn;
        fn = normal_force(kn, cn, depth, n, vn);
        if (law == BTSDBF) fs = smul(-ks * dt, vs) - smul(cs, vs);
        else fs = shear_rotation(h.s, vs) - smul(ks * dt, vs);
        double fmax = fc * kn * depth;
        if (bs == 0.0) {
            if (len(fs) > fmax) fs = smul(fmax, dir(fs));
            else fs = fs - smul(cs, vs);
        } else {
            if (len(fs) > bs) fs = smul(fmax, dir(fs));
            else fs = fs - smul(cs, vs);
        }
        if (law != BTSDBF) h.s = fs;
        h.f |= F_CONTACT;
    } else if (law != BSDB2F || (h.f & F_LASTBOND)) {
        d3 rv = va - vb;
        d3 vn = smul(dot(rv, n), n);
        if (law == BTSDBF) fn = normal_force(kn, cn, depth, n, vn);
        else fn = normal_force_bond(kn, cn, depth, n, vn);
        st_clean_contact(h);
    }
    d3 f = fn + fs;
    o.imp = muls(f, dt);
    o.ang_a = smul(dt, cross(ra, f));
    o.ang_b = smul(-dt, cross(rb, f));
    o.added = true;
    if (REC) {
        r.point = c.p; r.dir = c.n; r.overlap = c.depth;
        r.to_master = smul(-1.0, o.imp); r.ang_to_master = o.ang_b;
    }
}

}  // namespace vg
