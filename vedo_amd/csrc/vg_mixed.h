// vg_mixed.h -- the FP32/FP64 mixed-precision pair solve.
//
// Same physics and branch structure as solve() in vg_device.h (the five
// ImpactSolvers of the reference), with this split of precisions:
//   FP64: every decision that sets an ImpactStatus flag -- the overlap
//         depth of the contact detection (so the active set is the FP64
//         one, bit for bit), the normal-bond test -- the contact normal and
//         lever arms, the relative velocity at the contact point, the
//         normal spring and dashpot force (formed as the FP64 path forms
//         it: in a compressed bed a sphere's net impulse is a small
//         remainder of its contacts' normal forces, which FP32 rounding of
//         those forces would swamp), the stored history (shear force, Kn /
//         initial velocity) and its update, the impulse sums and the
//         integration (in the step kernel);
//   FP32: the tangential spring increment, the friction cap and the
//         tangential dashpot (the torques are formed from that tangential
//         force in FP64).
// The per-pair force error against the FP64 path is a few FP32 rounding
// units of the pair's tangential force.
#pragma once

#include "vg_device.h"

namespace vg {

struct f3 {
    float x, y, z;
};

__device__ inline f3 mkf(float x, float y, float z) { return f3{x, y, z}; }
__device__ inline f3 tof(d3 a) { return mkf((float)a.x, (float)a.y, (float)a.z); }
__device__ inline d3 tod(f3 a) { return mk((double)a.x, (double)a.y, (double)a.z); }
__device__ inline f3 operator+(f3 a, f3 b) { return mkf(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline f3 operator-(f3 a, f3 b) { return mkf(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline f3 operator-(f3 a) { return mkf(-a.x, -a.y, -a.z); }
__device__ inline f3 operator*(float s, f3 v) { return mkf(s * v.x, s * v.y, s * v.z); }
__device__ inline float dotf(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ inline f3 crossf(f3 a, f3 b)
{
    return mkf(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline float lenf(f3 a) { return sqrtf(dotf(a, a)); }
// unit vector; zero for the zero vector (Vector3d::direction)
__device__ inline f3 dirf(f3 a)
{
    const float l = lenf(a);
    return l != 0.0f ? (1.0f / l) * a : mkf(0.0f, 0.0f, 0.0f);
}

// Contact geometry of a pair for the mixed solve: the FP64 depth and
// activity, the FP32 normal (slave -> master) and lever arms from the slave
// and the master centres to the contact point.
struct ContactM {
    double depth;
    bool active;
    d3 n64;         // the normal
    d3 ra64, rb64;  // the lever arms from the slave and master centres
};

// CDSphere_Sphere (CDSphere2Sphere.cpp:14-60) for slave a, master b: the
// depth rA + rB - |xB - xA| in FP64 exactly as cd_ss forms it; the contact
// point P = xA + (rA - depth/2) n enters only through P - xA and P - xB.
__device__ inline void cd_ss_m(d3 xa, double ra, d3 xb, double rb, ContactM &c)
{
    const d3 im = xb - xa;
    const double l = len(im);
    c.depth = ra + rb - l;
    c.active = c.depth > 0;
    const double il = l != 0.0 ? 1.0 / l : 0.0;
    c.n64 = mk(im.x * il, im.y * il, im.z * il);
    c.ra64 = smul(ra - 0.5 * c.depth, c.n64);
    c.rb64 = c.ra64 - im;
}

// The same under a periodic boundary: the centre-to-centre vector as its
// minimum image (CDSphere2Sphere.cpp:19-22), so are both lever arms.
__device__ inline void cd_ss_m_bc(d3 xa, double ra, d3 xb, double rb, const PbcDev *b, ContactM &c)
{
    const d3 im = bc_diff(b, xb - xa);
    const double l = len(im);
    c.depth = ra + rb - l;
    c.active = c.depth > 0;
    const double il = l != 0.0 ? 1.0 / l : 0.0;
    c.n64 = mk(im.x * il, im.y * il, im.z * il);
    c.ra64 = smul(ra - 0.5 * c.depth, c.n64);
    c.rb64 = c.ra64 - im;
}

// A contact detected in FP64 (sphere-wall) in the mixed form; under a
// periodic boundary the lever arms are minimum images (ISwBSDBF.cpp:69-73).
__device__ inline void to_mixed(const Contact &c, d3 xa, d3 xb, ContactM &m, const PbcDev *b = nullptr)
{
    m.depth = c.depth;
    m.active = c.active;
    m.n64 = c.n;
    m.ra64 = bc_diff(b, c.p - xa);
    m.rb64 = bc_diff(b, c.p - xb);
}

// Relative velocity at the contact point, v = vA - vB + wA x rA - wB x rB,
// split into its normal part vn = (v.n) n and tangential part vs = v - vn,
// in FP64 (vn64, the normal part, stays FP64 for the normal force; vs is
// rounded to FP32).  The terms cancel: in a
// rolling contact the two spheres' rotational parts nearly annul each other,
// and in a resting one the centre velocities do, so FP32 terms would leave
// the small tangential velocity (whose direction sets a sliding contact's
// friction) with a relative error far above FP32's.
__device__ inline void rel_vel(d3 va, d3 vb, d3 wa, d3 wb, const ContactM &c, d3 &vn64, f3 &vs)
{
    const d3 v = ((va - vb) + cross(wa, c.ra64)) - cross(wb, c.rb64);
    vn64 = smul(dot(v, c.n64), c.n64);
    vs = tof(v - vn64);
}

// ShearForceRotation (ISwModels.cpp:110-123) of the FP64 history onto the
// FP32 direction of the tangential velocity, in FP64.
__device__ inline d3 shear_rotation_m(d3 hs, f3 vs)
{
    const float mag = sqrtf((float)dot(hs, hs));
    const f3 u = dirf(vs);
    const float sgn = (dot(hs, tod(vs)) <= 0.0) ? -mag : mag;
    return tod(sgn * u);
}

// Torques: a sphere's lever arm to the contact point lies along the contact
// normal, as do the normal spring and dashpot forces, so their torque is zero
// and the sphere's torque is that of the tangential force alone; forming it
// from the tangential force avoids the FP32 cancellation of ra x fn.  A
// wall's lever arm (master with zero sudo-mass) takes the whole force.
// xa: the slave's centre, for the recorded contact point only.
template <bool REC, int LS>
__device__ inline void solve_mixed(const LawDev &L, double ma, double mb, d3 xa, d3 va, d3 wa, d3 vb, d3 wb,
                                   const ContactM &c, double dt, Hist &h, PairOut &o, RecOut &r)
{
    o.added = false;
    o.linear_only = false;
    const int law = L.law;
    const double depth = c.depth;
    const float df = (float)depth, dtf = (float)dt;
    if (LS >= LS_ALL && law == LAW_HERTZ) {   // ISwHertz (ISwHertz.cpp:20-94)
        o.linear_only = true;
        if (!c.active) {
            st_clean(h);
            return;
        }
        d3 fn = mk(0, 0, 0);
        if (depth > 0.0) {
            d3 vn;
            f3 vs;
            rel_vel(va, vb, wa, wb, c, vn, vs);
            if (!(h.f & F_CONTACT)) {
                h.f |= F_CONTACT;
                h.aux = len(vn);
            }
            fn = smul(-L.p[0] * pow(depth, 1.5), c.n64) - smul(L.p[1], vn);
        } else {
            st_clean_contact(h);
        }
        o.imp = muls(fn, dt);
        o.ang_a = o.ang_b = mk(0, 0, 0);
        o.added = true;
        if (REC) {
            r.point = xa + c.ra64; r.dir = c.n64; r.overlap = depth;
            r.to_master = -o.imp; r.ang_to_master = mk(0, 0, 0);
        }
        return;
    }
    if (LS != LS_SD && law == LAW_H) {   // ISwH (ISwH.cpp:30-156)
        const double knh = L.p[0], meff = L.p[1], betan = L.p[4];
        if (!c.active) {
            st_clean(h);
            return;
        }
        f3 fst = mkf(0.0f, 0.0f, 0.0f);
        d3 f = mk(0, 0, 0), ra = mk(0, 0, 0), rb = ra;
        if (depth > 0.0) {
            ra = c.ra64;
            rb = c.rb64;
            d3 vn;
            f3 vs;
            rel_vel(va, vb, wa, wb, c, vn, vs);
            double kn;
            if (!(h.f & F_CONTACT)) {
                // first contact: Kn from the approach speed, kept in the history (FP64)
                const double vl = len(vn);
                if (vl == 0.0) kn = knh;
                else kn = L.p[6] * pow(vl, 0.4);   // p[6]: the law's factor, set on the host
                h.aux = kn;
                h.f |= F_CONTACT;
            } else {
                kn = h.aux;
            }
            const double ncd = 2.0 * sqrt(meff * kn);
            const d3 fn = normal_force(kn, ncd * betan, depth, c.n64, vn);
            const float knf = (float)kn;
            const float ks = (float)L.p[2] * cbrtf((float)L.p[3] * fabsf(knf * df));
            const float ff = (float)L.p[5] * knf * df;
            d3 fs = shear_rotation_m(h.s, vs) - tod((ks * dtf) * vs);
            const f3 fsf = tof(fs);
            if (lenf(fsf) >= ff) fs = tod(ff * dirf(fsf));
            h.s = fs;
            fst = tof(fs);
            f = fn + tod(fst);
        } else {
            st_clean_contact(h);
        }
        o.imp = muls(f, dt);
        o.ang_a = smul(dt, cross(ra, tod(fst)));
        o.ang_b = smul(-dt, cross(rb, mb == 0 ? f : tod(fst)));
        o.added = true;
        if (REC) {
            r.point = xa + c.ra64; r.dir = c.n64; r.overlap = depth;
            r.to_master = smul(-1.0, o.imp); r.ang_to_master = o.ang_b;
        }
        return;
    }
    if constexpr (LS == LS_H) return;
    // spring-dashpot family (ISwBSDBF.cpp:26-139, ISwBSDB2F.cpp:26-149, ISwBtSDBF.cpp:26-137)
    const double kn = L.p[0], bn = L.p[4], bs = L.p[5];
    const float knf = (float)kn, ksf = (float)L.p[1], fc = (float)L.p[6];
    float me;   // ISwModels::CriticalDamping (ISwModels.cpp:88-108)
    if (mb == 0) me = (float)ma;
    else if (ma == 0) me = (float)mb;
    else me = (float)(0.5 * (ma + mb));
    double me64;   // ISwModels::CriticalDamping, as the FP64 path forms it
    if (mb == 0) me64 = ma;
    else if (ma == 0) me64 = mb;
    else me64 = 0.5 * (ma + mb);
    const double cn = L.p[2] * (2.0 * sqrt(me64 * kn));
    const float cs = (float)L.p[3] * (2.0f * sqrtf(me * ksf));
    const bool nbond = (depth <= 0.0) && (kn != 0.0) && (bn != 0.0) && (-depth * kn < bn);
    if (!c.active && !nbond) {
        st_clean(h);
        return;
    }
    f3 fsf = mkf(0.0f, 0.0f, 0.0f);
    d3 fn = mk(0, 0, 0), ra = fn, rb = fn;
    if (depth > 0.0) {
        ra = c.ra64;
        rb = c.rb64;
        d3 vn;
        f3 vs;
        rel_vel(va, vb, wa, wb, c, vn, vs);
        fn = normal_force(kn, cn, depth, c.n64, vn);
        const float fmax = fc * knf * df;
        if (law == BTSDBF) {
            fsf = (-ksf * dtf) * vs - cs * vs;
            if (bs == 0.0 ? lenf(fsf) > fmax : lenf(fsf) > (float)bs) fsf = fmax * dirf(fsf);
            else fsf = fsf - cs * vs;
        } else {
            d3 fs = shear_rotation_m(h.s, vs) - tod((ksf * dtf) * vs);
            const f3 ft = tof(fs);
            if (bs == 0.0 ? lenf(ft) > fmax : lenf(ft) > (float)bs) fs = tod(fmax * dirf(ft));
            else fs = fs - tod(cs * vs);
            h.s = fs;
            fsf = tof(fs);
        }
        h.f |= F_CONTACT;
    } else if (law != BSDB2F || (h.f & F_LASTBOND)) {
        // tension of a normal bond: relative velocity of the centres only
        const d3 vn = smul(dot(va - vb, c.n64), c.n64);
        if (law == BTSDBF) fn = normal_force(kn, cn, depth, c.n64, vn);
        else fn = normal_force_bond(kn, cn, depth, c.n64, vn);
        st_clean_contact(h);
    }
    const d3 f = fn + tod(fsf);
    o.imp = muls(f, dt);
    o.ang_a = smul(dt, cross(ra, tod(fsf)));
    o.ang_b = smul(-dt, cross(rb, mb == 0 ? f : tod(fsf)));
    o.added = true;
    if (REC) {
        r.point = xa + c.ra64; r.dir = c.n64; r.overlap = depth;
        r.to_master = smul(-1.0, o.imp); r.ang_to_master = o.ang_b;
    }
}

}  // namespace vg
