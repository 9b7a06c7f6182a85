// vg_kernels.hip -- gfx950 kernels of the DEM step.
//
// Per step one fused kernel (k_step) does, for every sphere: contact detection
// against its neighbour list, the impact solver of every pair (both sides of a
// sphere pair evaluate the pair identically), the field/external impulse and
// DOSphere::Response.  A workgroup stages its 256 spheres in LDS, settles the
// surely separated pairs per sphere, queues the rest and solves the queue one
// entry per thread, so that lanes stay busy whatever the spread of contact
// counts; each sphere then adds its entries' impulses in list order.  It reads state buffer A and writes buffer B, so no
// thread sees a half-updated neighbour.  Per-sphere impulse sums follow the
// neighbour-list order, which the rebuild lays down in the reference's
// IactPairTab order (LeapConsultant.h:207-444), so the FP64 sums round exactly
// as the reference's DiscreteObject::AddImpact sequence does.
//
// Every up_iact steps the rebuild kernels (k_bounds ... k_inner_build)
// restate LeapConsultant::Reset: a cell grid of size ZoneRange over the local
// DOs, candidate pairs within safe_i + safe_j in the 13-cell half shell, self
// cell and global DOs, and IactRecordTab's history carry for pairs active at
// the rebuild.
#include <hip/hip_runtime.h>

#include "vg_kernels.h"
#include "vg_mixed.h"

namespace vg {

__constant__ int NB13[13][3] = VG_NB13_INIT;

__device__ inline size_t eidx(int h, int k, int ki, int n, int i)
{
    return ((size_t)(h * ki + k)) * (size_t)n + (size_t)i;
}
// Streaming accesses: data only the owning sphere touches (its orientation,
// list entries, next-step state) bypass the caches' retention so that L2
// keeps the neighbour positions and velocities many threads gather (the
// next step reads the new state from memory in any case: it does not fit
// L2).  Level 0: plain accesses; 1: streaming loads and list stores; 2: also
// the next-step state; 3: also the contact histories.
#ifndef VG_NT
#define VG_NT 2
#endif
template <class T> __device__ inline T nt_load(const T *p)
{
    if (VG_NT >= 1) return __builtin_nontemporal_load(p);
    return *p;
}
template <class T> __device__ inline void nt_store(T v, T *p)
{
    if (VG_NT >= 1) __builtin_nontemporal_store(v, p);
    else *p = v;
}
typedef double v4d __attribute__((ext_vector_type(4)));
__device__ inline void nt_store4(double4 v, double4 *p)
{
    if (VG_NT >= 2) __builtin_nontemporal_store(*reinterpret_cast<const v4d *>(&v), reinterpret_cast<v4d *>(p));
    else *p = v;
}
template <class T> __device__ inline void st_store(T v, T *p)
{
    if (VG_NT >= 2) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// History in slot `sl` of sphere i: the shear force always, the fourth
// word only where a Hertz law may have set it (AUX; without it the word is
// read as zero and written as zero).
typedef double v2d __attribute__((ext_vector_type(2)));
__device__ inline double2 hist_ld(const double2 *p)
{
    if (VG_NT >= 3) {
        const v2d x = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(p));
        return make_double2(x.x, x.y);
    }
    return *p;
}
__device__ inline void hist_st(double2 v, double2 *p)
{
    if (VG_NT >= 3) __builtin_nontemporal_store(*reinterpret_cast<const v2d *>(&v), reinterpret_cast<v2d *>(p));
    else *p = v;
}
template <bool AUX = true>
__device__ inline void load_hist(const Ell &e, int sl, int n, int i, Hist &hs)
{
    const double2 *p = reinterpret_cast<const double2 *>(e.hd + (size_t)sl * (size_t)n + (size_t)i);
    const double2 u = hist_ld(p), v = hist_ld(p + 1);
    hs.s = mk(u.x, u.y, v.x);
    hs.aux = AUX ? v.y : 0.0;
}
template <bool AUX = true>
__device__ inline void store_hist(const Ell &e, int sl, int n, int i, const Hist &hs)
{
    double2 *p = reinterpret_cast<double2 *>(e.hd + (size_t)sl * (size_t)n + (size_t)i);
    hist_st(make_double2(hs.s.x, hs.s.y), p);
    hist_st(make_double2(hs.s.z, AUX ? hs.aux : 0.0), p + 1);
}
// A set of a sphere's list positions or history slots, 0 .. 64 W - 1: one
// 64-bit word per 64.  The words are never indexed by a run-time value, so
// they stay in registers.
template <int W>
struct Bits {
    unsigned long long w[W];
    __device__ Bits()
    {
#pragma unroll
        for (int q = 0; q < W; ++q) w[q] = 0ull;
    }
    __device__ void set(int k)
    {
        if constexpr (W == 1) {
            w[0] |= 1ull << k;   // k < 64
            return;
        }
#pragma unroll
        for (int q = 0; q < W; ++q)
            if ((k >> 6) == q) w[q] |= 1ull << (k & 63);
    }
    __device__ bool has(int k) const
    {
        bool r = false;
#pragma unroll
        for (int q = 0; q < W; ++q)
            if ((k >> 6) == q) r = (w[q] >> (k & 63)) & 1ull;
        return r;
    }
    __device__ bool any() const
    {
        unsigned long long o = 0ull;
#pragma unroll
        for (int q = 0; q < W; ++q) o |= w[q];
        return o != 0ull;
    }
    __device__ int count() const
    {
        int c = 0;
#pragma unroll
        for (int q = 0; q < W; ++q) c += __popcll(w[q]);
        return c;
    }
    // the lowest member, which leaves the set (the set is not empty)
    __device__ int pop_first()
    {
        if constexpr (W == 1) {
            const int k = __builtin_ctzll(w[0]);
            w[0] &= w[0] - 1ull;
            return k;
        }
        int k = 0;
        bool done = false;
#pragma unroll
        for (int q = 0; q < W; ++q)
            if (!done && w[q]) {
                k = 64 * q + __builtin_ctzll(w[q]);
                w[q] &= w[q] - 1ull;
                done = true;
            }
        return k;
    }
    // the highest member leaves the set (the set is not empty)
    __device__ void drop_last()
    {
        if constexpr (W == 1) {
            w[0] &= ~(1ull << (63 - __builtin_clzll(w[0])));
            return;
        }
        bool done = false;
#pragma unroll
        for (int q = W - 1; q >= 0; --q)
            if (!done && w[q]) {
                w[q] &= ~(1ull << (63 - __builtin_clzll(w[q])));
                done = true;
            }
    }
    __device__ void remove(const Bits &o)
    {
#pragma unroll
        for (int q = 0; q < W; ++q) w[q] &= ~o.w[q];
    }
};

// The lowest history slot not in `used`, which then holds it.  A sphere's
// slots in use never outnumber its list entries (at most KI_MAX < 64 W).
template <int W>
__device__ inline int take_slot(Bits<W> &used)
{
    if constexpr (W == 1) {
        const int z = __builtin_ctzll(~used.w[0]);
        used.w[0] |= 1ull << z;
        return z;
    }
    int z = 0;
    bool done = false;
#pragma unroll
    for (int q = 0; q < W; ++q)
        if (!done && ~used.w[q]) {
            z = 64 * q + __builtin_ctzll(~used.w[q]);
            used.w[q] |= 1ull << (z & 63);
            done = true;
        }
    return z;
}

// ---------------------------------------------------------------------------
// Step
// ---------------------------------------------------------------------------

// Rebuild sphere i's neighbour list from its candidate list: keep partners
// within `skin` of contact, carry flags and history slots of partners already
// listed (both lists are ordered subsequences of the candidate list).
template <bool EXT>
__device__ inline int relist(const StepArgs &a, int i, d3 xi, double ri, int h_old, int c_old, int h_new,
                             double skin)
{
    const unsigned e = a.o_cnt[i];
    int p = 0, q = 0;
    unsigned old_partner = c_old > 0 ? (a.ell.ent[eidx(h_old, 0, a.ki, a.ld, i)] & PARTNER_MASK) : 0xffffffffu;
    for (unsigned t = 0; t < e; ++t) {
        const unsigned j = a.o_nbr[(size_t)t * a.ld + i] & PARTNER_MASK;
        Contact c;
        bool near = true;
        if (j < (unsigned)a.ld) {
            const double4 pj = a.in.pos[j];
            d3 dx = ld3(pj) - xi;
            if constexpr (EXT) dx = bc_diff(a.pbc, dx);
            const double lim = ri + pj.w + skin;
            // far beyond the skin: excluded without the exact geometry
            if (dot(dx, dx) > lim * lim * (1.0 + 1e-9)) near = false;
            else if constexpr (EXT) cd_ss_bc(xi, ri, ld3(pj), pj.w, a.pbc, c);
            else cd_ss(xi, ri, ld3(pj), pj.w, c);
        } else {
            cd_wall_at<EXT>(xi, ri, a.walls[j - a.ld], a.holes, (int)(j - a.ld), c);
        }
        const bool carried = (p < c_old) && (old_partner == j);
        unsigned f = 0;
        unsigned char sl = 0;
        if (carried) {
            f = a.ell.ent[eidx(h_old, p, a.ki, a.ld, i)] & ~PARTNER_MASK;
            if (f & F_CONTACT) sl = a.ell.slot[eidx(h_old, p, a.ki, a.ld, i)];
            ++p;
            old_partner = p < c_old ? (a.ell.ent[eidx(h_old, p, a.ki, a.ld, i)] & PARTNER_MASK) : 0xffffffffu;
        }
        if (near && c.depth >= -skin) {
            if (q >= a.ki) {
                atomicOr(&a.flags[3], ERR_INNER_CAP);
                break;
            }
            if (!carried) {
                // A candidate outside the list has been solved as inactive
                // every step: its status is Clean() (ImpactStatus.cpp:259-274).
                f = 0;
                if (j >= (unsigned)a.ld || a.in.oid[i] < a.in.oid[j]) f |= F_SLAVE;
            }
            a.ell.ent[eidx(h_new, q, a.ki, a.ld, i)] = j | f;
            // a carried contact keeps its history where it is
            if (f & F_CONTACT) a.ell.slot[eidx(h_new, q, a.ki, a.ld, i)] = sl;
            ++q;
        }
    }
    return q;
}

// History words are meaningful only while the entry's Contact flag is set:
// every other status of the five solvers has a zero shear force and a zero
// Kn / initial velocity (Clean, CleanContactStatusAndImpact, ImpactStatus()).
// So the step reads them only for contacts and writes them only when the
// solve leaves the pair in contact.
// Workgroups are dealt to the 8 XCDs round-robin; remap them so that each XCD
// walks one contiguous eighth of the (spatially sorted) spheres and its L2
// holds the neighbours its waves gather.
__device__ inline int xcd_block(int b, int nb)
{
    const int per = (nb + 7) / 8, x = b % 8, q = b / 8;
    const int full = nb - 8 * (per - 1);   // XCDs 0..full-1 get `per` blocks, the rest per-1
    return x < full ? x * per + q : full * per + (x - full) * (per - 1) + q;
}

// One entry (owner sphere i, list half h, slot k) of the pair solves: the
// backup of the previous step's status, the contact detection and the law,
// the entry's new status and history, the constrained-wall share and the
// record.  Returns the owner's impulse and angular impulse from the pair
// (zero where the solver adds nothing) and whether the pair stays in contact
// on its slave side.
// `ent` is the entry's word as pass 1 read it (carried in the queue, so the
// partner's state is requested together with the history); `sl` its history
// slot, or the free slot it takes if the solve starts a contact.  UNI: one scene
// whose spheres share one group; its sphere-sphere law (a.law_ss) and scene
// row are read through uniform addresses.
template <bool REC, int LS, bool MIX, bool UNI>
__device__ inline bool solve_entry(const StepArgs &a, int i, int h, int k, unsigned ent, int sl, d3 xi, double ri,
                                   d3 vi, d3 wi, double mi, int gi, const SceneDev &S, const double4 *s_pos,
                                   const double4 *s_vel, const double4 *s_avl, unsigned b0, unsigned nb, d3 &dI,
                                   d3 &dL)
{
    dI = mk(0, 0, 0);
    dL = mk(0, 0, 0);
    const size_t ei = eidx(h, k, a.ki, a.ld, i);
    const unsigned j = ent & PARTNER_MASK;
    Hist hs;
    hs.f = ent & ~PARTNER_MASK;
    const bool had_contact = (hs.f & F_CONTACT) != 0;
    // IactContainer::BackupImpactStatus of the previous step (IactContainer.cpp:128-131)
    if (!(hs.f & F_NOBACKUP)) {
        unsigned f = hs.f & ~(F_LASTCONTACT | F_LASTBOND);
        if (hs.f & F_CONTACT) f |= F_LASTCONTACT;
        if (hs.f & F_BOND) f |= F_LASTBOND;
        hs.f = f;
    }
    hs.f &= ~F_NOBACKUP;
    // meaningful only for contacts (most queued entries are)
    if (had_contact) {
        load_hist<LS != LS_SD>(a.ell, sl, a.ld, i, hs);
    } else {
        hs.s = mk(0, 0, 0);
        hs.aux = 0.0;
    }
    PairOut o;
    o.added = false;
    RecOut r;
    if (REC) {
        r.point = r.dir = r.to_master = r.ang_to_master = mk(0, 0, 0);
        r.overlap = 0.0;
    }
    // The pair's slave (A) and master (B) inputs: the owner is A on its slave
    // side and against walls, B on its master side.  One contact detection
    // and one law evaluation serve every case, so a wavefront runs the solver
    // once per pass whatever mix of sides its lanes hold.
    Contact c;
    ContactM cm;
    d3 xa, va, wa, xb, vb, wb;
    double ma, mb;
    // the pair's IactModel: one-scene handles read its fields at once, from a
    // uniform address for sphere pairs; batches keep it in memory until the
    // solver reads each field (fewer registers in the larger kernels)
    [[maybe_unused]] LawDev lawv;
    const LawDev *lawp = nullptr;
    bool owner_a = true;
    const WallDev *w = nullptr;
    if (j < (unsigned)a.ld) {
        const bool loc = j - b0 < nb;
        const double4 pj = loc ? s_pos[j - b0] : a.in.pos[j];
        const double4 vj = loc ? s_vel[j - b0] : a.in.vel[j];
        const double4 wj = loc ? s_avl[j - b0] : a.in.avl[j];
        if constexpr (UNI) {
            lawv = a.t.laws[a.law_ss];
        } else {
            const int gj = a.single_group ? gi : a.t.models[a.in.mdl[j]].group;
            lawp = &a.t.laws[S.law_off + gi * S.n_groups + gj];
        }
        owner_a = (hs.f & F_SLAVE) != 0;
        const d3 xj = ld3(pj), vjj = ld3(vj), wjj = ld3(wj);
        xa = owner_a ? xi : xj;
        va = owner_a ? vi : vjj;
        wa = owner_a ? wi : wjj;
        ma = owner_a ? mi : wj.w;
        xb = owner_a ? xj : xi;
        vb = owner_a ? vjj : vi;
        wb = owner_a ? wjj : wi;
        mb = owner_a ? wj.w : mi;
        const double ra = owner_a ? ri : pj.w, rb = owner_a ? pj.w : ri;
        if constexpr (LS == LS_EXT) {
            if (MIX) cd_ss_m_bc(xa, ra, xb, rb, a.pbc, cm);
            else cd_ss_bc(xa, ra, xb, rb, a.pbc, c);
        } else {
            if (MIX) cd_ss_m(xa, ra, xb, rb, cm);
            else cd_ss(xa, ra, xb, rb, c);
        }
    } else {
        w = &a.walls[j - a.ld];
        if constexpr (UNI) lawv = a.t.laws[S.law_off + gi * S.n_groups + w->group];
        else lawp = &a.t.laws[S.law_off + gi * S.n_groups + w->group];
        cd_wall_at<LS == LS_EXT>(xi, ri, *w, a.holes, (int)(j - a.ld), c);
        xa = xi; va = vi; wa = wi; ma = mi;
        xb = w->pos; vb = w->vel; wb = w->av; mb = w->msudo;
        if constexpr (LS == LS_EXT) {
            if (MIX) to_mixed(c, xi, w->pos, cm, a.pbc);
        } else {
            if (MIX) to_mixed(c, xi, w->pos, cm);
        }
    }
    const LawDev &law = UNI ? lawv : *lawp;
    if (MIX) solve_mixed<REC, LS>(law, ma, mb, xa, va, wa, vb, wb, cm, S.dt, hs, o, r);
    else if constexpr (LS == LS_EXT) solve<REC, LS>(law, ma, mb, xa, va, wa, xb, vb, wb, c, S.dt, hs, o, r, a.pbc);
    else solve<REC, LS>(law, ma, mb, xa, va, wa, xb, vb, wb, c, S.dt, hs, o, r);
    if (o.added) {
        if (owner_a) {
            dI = o.imp;
            if (!o.linear_only) dL = o.ang_a;
        } else {
            dI = o.linear_only ? -o.imp : smul(-1.0, o.imp);
            if (!o.linear_only) dL = o.ang_b;
        }
        if (w && w->constrained) {
            // the wall's share (DiscreteObject::AddImpact on the master), added in pair order by k_cw_sum
            const int ws = (int)(j - a.ld);
            int p = -1;
            for (int q = 0; q < CW_SLOTS && p < 0; ++q)
                if (a.cw_w[(size_t)q * a.ld + i] == ws) p = a.cw_p[(size_t)q * a.ld + i];
            if (p >= 0) {
                const d3 m = o.linear_only ? -o.imp : smul(-1.0, o.imp);
                double *v = a.cw_val + 3 * (size_t)p;
                v[0] = m.x; v[1] = m.y; v[2] = m.z;
                a.cw_stamp[p] = a.stamp;
            } else {
                atomicOr(&a.flags[3], ERR_CW);
            }
        }
    }
    if (hs.f != (ent & ~PARTNER_MASK)) nt_store(j | hs.f, a.ell.ent + ei);
    bool slave_contact = false;
    if (hs.f & F_CONTACT) {
        store_hist<LS != LS_SD>(a.ell, sl, a.ld, i, hs);
        if (!had_contact) a.ell.slot[ei] = (unsigned char)sl;
        slave_contact = (hs.f & F_SLAVE) != 0;
    }
    // Without bonds a status without Contact is Clean (all record fields
    // zero): only contacts store their record, readers take the rest as zero
    if (REC && ((hs.f & F_CONTACT) || !a.no_bond)) {
        double *rp = a.rec + ((size_t)k * 13) * a.ld + i;
        const double vals[13] = {r.point.x, r.point.y, r.point.z, r.dir.x, r.dir.y, r.dir.z,
                                 r.to_master.x, r.to_master.y, r.to_master.z, r.ang_to_master.x,
                                 r.ang_to_master.y, r.ang_to_master.z, r.overlap};
        for (int q = 0; q < 13; ++q) rp[(size_t)q * a.ld] = vals[q];
    }
    return slave_contact;
}

// Exclusive prefix sum of v over the block's NT threads; *total gets the sum.
template <int NT>
__device__ inline int block_scan(int v, int *s_wsum, int &total)
{
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_wsum[w] = x;
    __syncthreads();
    int base = 0;
    for (int q = 0; q < w; ++q) base += s_wsum[q];
    total = 0;
#pragma unroll
    for (int q = 0; q < NT / 64; ++q) total += s_wsum[q];
    __syncthreads();
    return base + x - v;
}

// 3 workgroups per CU (12 wavefronts), as the LDS allows: the register
// budget is capped at 168 VGPRs, the larger law sets spill a few registers
template <bool REC, int LS, bool MIX, bool UNI, int W>
__global__ __launch_bounds__(STEP_NT, STEP_WG_PER_CU) void k_step(StepArgs a)
{
    constexpr int NT = STEP_NT;   // threads per block: the launch's block size
    constexpr int QPT = 8;    // queued entries per thread and round of the queue
    // the block's spheres (partners inside the block are read from here) and
    // its queue of pair solves (item, entry word, history slot), one per
    // thread per pass, with their results; pass 1 stages the words and slots
    // of each thread's first QPT queued entries in the result area
    __shared__ double4 s_pos[NT], s_vel[NT], s_avl[NT];
    __shared__ int s_mdl[NT];
    // queue item: thread << (QB + 1) | list half << QB | entry, QB bits for the entry
    constexpr int QB = W == 1 ? 6 : 7;
    static_assert((NT - 1) << (QB + 1) < 65536, "a queue item fits 16 bits");
    __shared__ unsigned short s_item[NT * QPT];
    __shared__ unsigned s_qent[NT * QPT];
    __shared__ unsigned char s_qslot[NT * QPT];
    __shared__ double s_out[6][NT];
    __shared__ int s_wsum[NT / 64];
    __shared__ unsigned s_ncontact;
    static_assert(QPT * NT * 5 <= sizeof(s_out), "the staging fits the result area");
    unsigned *const s_stage = reinterpret_cast<unsigned *>(&s_out[0][0]);                   // [QPT][NT]
    unsigned char *const s_stage_slot = reinterpret_cast<unsigned char *>(s_stage + QPT * NT);   // [QPT][NT]

    const int t = threadIdx.x;
    const int b0 = xcd_block(blockIdx.x, gridDim.x) * NT;
    const int i = b0 + t;
    if (i < a.n_scenes) {
        // scene i's relist ring: clear the slot of the step after next, count this step's relist
        a.rflags[3 * i + (a.step + 2) % 3] = 0;
        if (a.rflags[3 * i + a.step % 3]) atomicAdd(a.relists, 1ull);
    }
    if (b0 >= a.n) return;   // the whole block is past the spheres (uniform)
    const unsigned nb = (unsigned)min(NT, a.n - b0);
    const bool in = i < a.n;
    const int mdl_i = in ? a.in.mdl[i] : 0;
    double4 p4 = make_double4(0.0, 0.0, 0.0, 0.0), v4 = p4, w4 = p4;
    if (in) {
        p4 = a.in.pos[i];
        v4 = a.in.vel[i];
        w4 = a.in.avl[i];
    }
    s_pos[t] = p4;
    s_vel[t] = v4;
    s_avl[t] = w4;
    s_mdl[t] = mdl_i;
    if (t == 0) s_ncontact = 0;
    // UNI: the scene is slot 0 (its row is read through uniform addresses)
    const int sc = UNI ? 0 : a.t.models[mdl_i].scene;
    const SceneDev &S = a.scenes[sc];
    const bool live = in && S.active > 0;
    if (in && !live) {
        // a frozen scene (finished, awaiting retirement) is carried over unchanged
        a.out.pos[i] = p4;
        a.out.vel[i] = v4;
        a.out.avl[i] = w4;
        for (int c = 0; c < 6; ++c) a.out.ori[(size_t)c * a.ld + i] = a.in.ori[(size_t)c * a.ld + i];
    }
    const d3 xi = ld3(p4);
    const double ri = p4.w;

    int h = 0, cnt = 0;
    bool relisted = false;
    if (live) {
        const bool do_relist = a.rflags[3 * sc + a.step % 3] != 0;
        const unsigned hdr = a.ell.hdr[i];
        h = hdr >> 7;
        cnt = hdr & 127;
        if (do_relist) {
            const int hn = h ^ 1;
            cnt = relist<LS == LS_EXT>(a, i, xi, ri, h, cnt, hn, S.skin);
            h = hn;
            a.ell.hdr[i] = (unsigned char)((h << 7) | cnt);
            relisted = true;
        }
    }
    __syncthreads();

    // Pass 1: surely separated sphere pairs are settled here (Clean status,
    // no impulse); every other entry is queued for the full solve.
    // W 64-bit words per set: lists of up to 64 entries take one
    Bits<W> todo;
    Bits<W> used;   // history slots of the entries in contact
    int nstage = 0;
    // entries PE at a time (their words and slots in one round of loads),
    // then in groups of PF: the partner-position gathers of a group are
    // independent and overlap in flight
    constexpr int PE = 12, PF = 6;
    for (int k1 = 0; k1 < cnt; k1 += PE) {
        unsigned ew[PE];
        unsigned char es[PE];
#pragma unroll
        for (int u = 0; u < PE; ++u) {
            ew[u] = k1 + u < cnt ? nt_load(a.ell.ent + eidx(h, k1 + u, a.ki, a.ld, i)) : PARTNER_MASK;
            es[u] = k1 + u < cnt ? nt_load(a.ell.slot + eidx(h, k1 + u, a.ki, a.ld, i)) : 0;
        }
#pragma unroll
        for (int g = 0; g < PE; g += PF) {
            if (k1 + g >= cnt) break;
            double4 pj[PF];
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const unsigned j = ew[g + u] & PARTNER_MASK;
                pj[u] = (a.no_bond && j < (unsigned)a.ld) ? (j - (unsigned)b0 < nb ? s_pos[j - b0] : a.in.pos[j])
                                                           : make_double4(0.0, 0.0, 0.0, -1.0);
            }
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int k = k1 + g + u;
                if (k >= cnt) break;
                const unsigned e = ew[g + u], j = e & PARTNER_MASK;
                if (e & F_CONTACT) used.set(es[g + u]);
                bool full = true;
                if (pj[u].w >= 0.0) {   // a sphere partner, unbonded laws
                    d3 dx = ld3(pj[u]) - xi;
                    if constexpr (LS == LS_EXT) dx = bc_diff(a.pbc, dx);
                    const double rs = ri + pj[u].w;
                    if (dot(dx, dx) > rs * rs * (1.0 + 1e-9)) {
                        full = false;
                        // Backup, then Clean() of an inactive pair (ImpactStatus.cpp:259-274)
                        // (its record is all zero: readers take a record without
                        // Contact, in a scene without bonds, as zero)
                        const unsigned f = e & ~PARTNER_MASK & ~F_STATUS & ~F_NOBACKUP;
                        if (f != (e & ~PARTNER_MASK)) nt_store(j | f, a.ell.ent + eidx(h, k, a.ki, a.ld, i));
                    }
                }
                if (full) {
                    todo.set(k);
                    if (nstage < QPT) {
                        s_stage[nstage * NT + t] = e;
                        s_stage_slot[nstage * NT + t] = es[g + u];
                    }
                    ++nstage;
                }
            }
        }
    }

    // Pass 2: the block's queued entries are solved one per thread per pass
    // (each thread takes at most QPT of its own entries per round, in list
    // order); after each pass every owner adds its entries' impulses in list
    // order, so the sums round exactly as a sequential walk of its list.
    d3 I = mk(0, 0, 0), L = mk(0, 0, 0);
    unsigned n_contact = 0;
    Bits<W> rem = todo;
    bool first_round = true;
    for (;;) {
        Bits<W> take = rem;
        for (int q = take.count(); q > QPT; --q) take.drop_last();
        const int c = take.count();
        int total = 0;
        const int off = block_scan<NT>(c, s_wsum, total);
        {
            // the first round takes the thread's first queued entries, whose
            // words and slots pass 1 staged in order; later rounds read them
            // again.  An entry not in contact gets a free slot for the
            // history it keeps if the solve starts a contact.
            Bits<W> tk = take;
            for (int q = 0; q < c; ++q) {
                const int k = tk.pop_first();
                const size_t x = eidx(h, k, a.ki, a.ld, i);
                const unsigned w = first_round ? s_stage[q * NT + t] : a.ell.ent[x];
                int sl;
                if (w & F_CONTACT) sl = first_round ? s_stage_slot[q * NT + t] : a.ell.slot[x];
                else sl = take_slot(used);
                s_item[off + q] = (unsigned short)((t << (QB + 1)) | (h << QB) | k);
                s_qent[off + q] = w;
                s_qslot[off + q] = (unsigned char)sl;
            }
        }
        rem.remove(take);
        first_round = false;
        __syncthreads();
        for (int r0 = 0; r0 < total; r0 += NT) {
            const int q = r0 + t;
            if (q < total) {
                const unsigned it = s_item[q];
                const int lt = (int)(it >> (QB + 1)), hq = (int)((it >> QB) & 1), k = (int)(it & ((1 << QB) - 1));
                const int io = b0 + lt;
                const double4 po = s_pos[lt], vo = s_vel[lt], wo = s_avl[lt];
                int go;
                const SceneDev *So;
                if (UNI) {
                    go = a.group0;
                    So = &a.scenes[0];
                } else {
                    const ModelDev &mo = a.t.models[s_mdl[lt]];
                    go = mo.group;
                    So = &a.scenes[mo.scene];
                }
                d3 dI, dL;
                const bool sc_ = solve_entry<REC, LS, MIX, UNI>(a, io, hq, k, s_qent[q], s_qslot[q], ld3(po), po.w,
                                                                ld3(vo), ld3(wo), wo.w, go, *So, s_pos, s_vel,
                                                                s_avl, (unsigned)b0, nb, dI, dL);
                n_contact += sc_ ? 1u : 0u;
                s_out[0][t] = dI.x; s_out[1][t] = dI.y; s_out[2][t] = dI.z;
                s_out[3][t] = dL.x; s_out[4][t] = dL.y; s_out[5][t] = dL.z;
            }
            __syncthreads();
            const int lo = max(off, r0), hi = min(off + c, r0 + NT);
            for (int q2 = lo; q2 < hi; ++q2) {
                const int u = q2 - r0;
                I = I + mk(s_out[0][u], s_out[1][u], s_out[2][u]);
                L = L + mk(s_out[3][u], s_out[4][u], s_out[5][u]);
            }
            __syncthreads();
        }
        if (!__syncthreads_or(rem.any())) break;
    }
    if (n_contact) atomicAdd(&s_ncontact, n_contact);

    if (live) {
        // the sphere's own state again from LDS (not held in registers through the solves)
        const double4 p4b = s_pos[t], v4b = s_vel[t], w4b = s_avl[t];
        const d3 xi = ld3(p4b), vi = ld3(v4b), wi = ld3(w4b);
        const double ri = p4b.w, mi = w4b.w;
        const int sc = UNI ? 0 : a.t.models[s_mdl[t]].scene;
        const SceneDev &S = a.scenes[sc];
        // DOContainer::AddFieldImpact + AddConstrainedImpact (DOContainer.cpp:20-50).
        // Where no model has an external force the second term adds +0 to a
        // sum that is never -0 (it starts at +0), which leaves it unchanged.
        const d3 gdt = muls(mk(S.gx, S.gy, S.gz), S.dt);
        I = I + muls(gdt, mi);
        if (!UNI || a.any_fext) {
            const ModelDev &md = a.t.models[s_mdl[t]];
            I = I + smul(S.dt, mk(md.fext[0], md.fext[1], md.fext[2]));
        }
        // DOSphere::Response (DOSphere.cpp:21-48)
        const d3 dv = smul(1.0 / mi, I);
        const d3 dav = smul(2.5 / mi / ri / ri, L);
        const d3 dp = smul(S.dt, vi + smul(0.5, dv));
        const d3 dw = smul(S.dt, wi + smul(0.5, dav));
        d3 xn = xi + dp;
        // DOContainer::EnforcePeriodicBoundaryConditions (SimMediator.cpp:746)
        if constexpr (LS == LS_EXT) xn = bc_enforce(a.pbc, xn);
        const double *oi = a.in.ori + i;
        const size_t LD = (size_t)a.ld;
        d3 oxn = mk(nt_load(oi), nt_load(oi + LD), nt_load(oi + 2 * LD));
        d3 ozn = mk(nt_load(oi + 3 * LD), nt_load(oi + 4 * LD), nt_load(oi + 5 * LD));
        rotate_pair(oxn, ozn, dw);
        const d3 vn = vi + dv, wn = wi + dav;
        const double path = (relisted ? 0.0 : v4b.w) + len(dp);
        if (path > S.relist_half) atomicOr(&a.rflags[3 * sc + (a.step + 1) % 3], 1);
        nt_store4(make_double4(xn.x, xn.y, xn.z, ri), a.out.pos + i);
        nt_store4(make_double4(vn.x, vn.y, vn.z, path), a.out.vel + i);
        nt_store4(make_double4(wn.x, wn.y, wn.z, mi), a.out.avl + i);
        double *oo = a.out.ori + i;
        st_store(oxn.x, oo); st_store(oxn.y, oo + LD); st_store(oxn.z, oo + 2 * LD);
        st_store(ozn.x, oo + 3 * LD); st_store(ozn.y, oo + 4 * LD); st_store(ozn.z, oo + 5 * LD);
        if (REC) {
            a.rec_imp[i] = I.x; a.rec_imp[a.ld + i] = I.y; a.rec_imp[2 * a.ld + i] = I.z;
            a.rec_imp[3 * a.ld + i] = L.x; a.rec_imp[4 * a.ld + i] = L.y; a.rec_imp[5 * a.ld + i] = L.z;
        }
    }
    // the block's count of pairs left in contact (slave sides) goes to its first wavefront's slot
    __syncthreads();
    if (t == 0 && s_ncontact) a.contacts[b0 >> 6] += s_ncontact;
}

template __global__ void k_step<false, LS_SD, false, false>(StepArgs);
template __global__ void k_step<true, LS_SD, false, false>(StepArgs);
template __global__ void k_step<false, LS_H, false, false>(StepArgs);
template __global__ void k_step<true, LS_H, false, false>(StepArgs);
template __global__ void k_step<false, LS_ALL, false, false>(StepArgs);
template __global__ void k_step<true, LS_ALL, false, false>(StepArgs);
template __global__ void k_step<false, LS_SD, true, false>(StepArgs);
template __global__ void k_step<true, LS_SD, true, false>(StepArgs);
template __global__ void k_step<false, LS_H, true, false>(StepArgs);
template __global__ void k_step<true, LS_H, true, false>(StepArgs);
template __global__ void k_step<false, LS_ALL, true, false>(StepArgs);
template __global__ void k_step<true, LS_ALL, true, false>(StepArgs);
template __global__ void k_step<false, LS_EXT, false, false>(StepArgs);
template __global__ void k_step<true, LS_EXT, false, false>(StepArgs);
template __global__ void k_step<false, LS_EXT, true, false>(StepArgs);
template __global__ void k_step<true, LS_EXT, true, false>(StepArgs);
template __global__ void k_step<false, LS_SD, false, true>(StepArgs);
template __global__ void k_step<true, LS_SD, false, true>(StepArgs);
template __global__ void k_step<false, LS_H, false, true>(StepArgs);
template __global__ void k_step<true, LS_H, false, true>(StepArgs);
template __global__ void k_step<false, LS_ALL, false, true>(StepArgs);
template __global__ void k_step<true, LS_ALL, false, true>(StepArgs);
template __global__ void k_step<false, LS_SD, true, true>(StepArgs);
template __global__ void k_step<true, LS_SD, true, true>(StepArgs);
template __global__ void k_step<false, LS_H, true, true>(StepArgs);
template __global__ void k_step<true, LS_H, true, true>(StepArgs);
template __global__ void k_step<false, LS_ALL, true, true>(StepArgs);
template __global__ void k_step<true, LS_ALL, true, true>(StepArgs);
// wide lists (more than 64 entries per sphere): the general law sets only
template __global__ void k_step<false, LS_ALL, false, false, 2>(StepArgs);
template __global__ void k_step<true, LS_ALL, false, false, 2>(StepArgs);
template __global__ void k_step<false, LS_ALL, true, false, 2>(StepArgs);
template __global__ void k_step<true, LS_ALL, true, false, 2>(StepArgs);
template __global__ void k_step<false, LS_ALL, false, true, 2>(StepArgs);
template __global__ void k_step<true, LS_ALL, false, true, 2>(StepArgs);
template __global__ void k_step<false, LS_ALL, true, true, 2>(StepArgs);
template __global__ void k_step<true, LS_ALL, true, true, 2>(StepArgs);
template __global__ void k_step<false, LS_EXT, false, false, 2>(StepArgs);
template __global__ void k_step<true, LS_EXT, false, false, 2>(StepArgs);
template __global__ void k_step<false, LS_EXT, true, false, 2>(StepArgs);
template __global__ void k_step<true, LS_EXT, true, false, 2>(StepArgs);

// Wall responses after the spheres' (DOFixedQuasiPlate.h:16-19,
// DOOrbitalQuasiPlate.cpp:18-34) and kinematic prescriptions
// (SimMediator.cpp:1121-1163).  kin 1: host-assigned status in `assigned`;
// kin 2: rigid rotation about (c, a) at omega from the reference status.
__global__ void k_walls(WallDev *walls, int nw, const SceneDev *scenes, double *stime, const double *assigned,
                        const WallDev *ref, const WallKin *kin, const double *cw_imp, int *rflags, int step)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const WallDev old = walls[w];
    if (old.scene < 0) return;   // slot not in use
    const SceneDev &S = scenes[old.scene];
    // scene time before / after this step, double-buffered by step parity
    // (SystemParameter::NextStep's tCurrent += dt); the scene's first wall
    // writes the next value
    const int sc = old.scene, nsc_t = step & 1;
    const double t_old = stime[2 * sc + nsc_t];
    if (S.active <= 0) {
        if (w == S.w_begin) stime[2 * sc + (nsc_t ^ 1)] = t_old;
        return;
    }
    const double dt = S.dt, t_new = t_old + dt;
    if (w == S.w_begin) stime[2 * sc + (nsc_t ^ 1)] = t_new;
    WallDev W = old;
    if (W.kin == 1) {
        const double *s = assigned + 15 * w;
        W.pos = mk(s[0], s[1], s[2]); W.vel = mk(s[3], s[4], s[5]);
        W.ox = mk(s[6], s[7], s[8]); W.oz = mk(s[9], s[10], s[11]); W.av = mk(s[12], s[13], s[14]);
    } else if (W.kin == 2) {
        const WallKin &K = kin[w];
        const WallDev &R = ref[w];
        const double th = K.omega * (t_new - K.t_ref);
        const double ct = cos(th), st = sin(th);
        auto rot = [&](d3 v) { return (smul(ct, v) + smul(st, cross(K.a, v))) + smul(dot(v, K.a) * (1.0 - ct), K.a); };
        W.pos = K.c + rot(R.pos - K.c);
        const d3 om = smul(K.omega, K.a);
        W.vel = cross(om, W.pos - K.c);
        W.ox = rot(R.ox);
        W.oz = rot(R.oz);
        W.av = om;
    } else if (W.behavior == 4) {   // orbital
        const d3 dp = muls(W.vel, dt), dw = muls(W.av, dt);
        W.pos = W.pos + dp;
        W.ox = rotate_around(W.ox, dw);
        W.oz = rotate_around(W.oz, dw);
    } else if (W.constrained) {
        // DOConstrainedQuasiPlate/Cylinder::Response (DOConstrainedQuasiPlate.cpp:18-35):
        // the impulse's component along Oz accelerates the wall; it turns at its own rate
        const d3 I = mk(cw_imp[3 * (size_t)w], cw_imp[3 * (size_t)w + 1], cw_imp[3 * (size_t)w + 2]);
        const d3 V = W.vel, Oz = W.oz;
        const d3 dv = smul(1.0 / W.msudo, smul(dot(I, Oz), dir(Oz)));
        const d3 dp = smul(dt, V + smul(0.5, dv));
        const d3 dw = muls(W.av, dt);
        W.pos = W.pos + dp;
        W.ox = rotate_around(W.ox, dw);
        W.oz = rotate_around(Oz, dw);
        W.vel = V + dv;
    }
    // Relist guard for moving walls.  A sphere-wall depth is the sphere's
    // distance to the plate's mid-plane rectangle (the cylinder's axis
    // segment) subtracted from constants (cd_sp, cd_sc), so it changes by at
    // most the largest displacement of a point of that rectangle (segment).
    // Under a rigid motion a point's displacement is affine in its position
    // on the rectangle, so the largest is at a corner (an end); the bound
    // allows for the rounding of the corners.  Accumulated since the last
    // relist, together with a sphere's own path it bounds every depth change:
    // a wall whose bound exceeds half the skin makes its scene's spheres
    // relist at the next step, as a sphere's own path does.  The bound
    // restarts when the step that just ran relisted (the lists were made at
    // these positions).
    double d = 0.0;
    if (W.shape == SH_PLATE) {
        const d3 oy0 = cross(old.oz, old.ox), oy1 = cross(W.oz, W.ox);
        const double hw = 0.5 * W.a0, hl = 0.5 * W.a2;
        for (int c = 0; c < 4; ++c) {
            const double sx = (c & 1) ? hw : -hw, sy = (c & 2) ? hl : -hl;
            const d3 c0 = (old.pos + muls(old.ox, sx)) + muls(oy0, sy);
            const d3 c1 = (W.pos + muls(W.ox, sx)) + muls(oy1, sy);
            d = fmax(d, len(c1 - c0));
        }
    } else {
        const double hh = 0.5 * W.a1;
        for (int c = 0; c < 2; ++c) {
            const double sz = c ? hh : -hh;
            d = fmax(d, len((W.pos + muls(W.oz, sz)) - (old.pos + muls(old.oz, sz))));
        }
    }
    d = d * (1.0 + 1e-9) + 1e-15 * (len(W.pos) + W.reach);
    W.path = (rflags[3 * sc + step % 3] ? 0.0 : old.path) + d;
    if (W.path > S.relist_half) atomicOr(&rflags[3 * sc + (step + 1) % 3], 1);
    walls[w] = W;
}

// Fresh lists after a rebuild: no scene relists at the next step (its ring
// slot r) ...
__global__ void k_ring_clear(int *rflags, int scap, int r)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < scap) rflags[3 * k + r] = 0;
}

// ... and the wall displacement bounds restart.
__global__ void k_wall_path_reset(WallDev *walls, int nw)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nw) walls[w].path = 0.0;
}

// ---------------------------------------------------------------------------
// Rebuild
// ---------------------------------------------------------------------------

__device__ inline unsigned long long ord64(double d)
{
    unsigned long long u = __double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// Bounds of each scene's local DOs (LeapConsultant.cpp:119-164): max / min
// position, max |v|, max range, as order-preserving integers reduced with
// atomicMax into red[scene][8]; a block whose spheres all belong to one scene
// (the usual case) reduces in LDS first.  Also the rank -> current-slot map
// of the next sort and the coordinates for the quantile sorts.
__global__ void k_bounds(Geo g, int n, const SceneDev *scenes, const ModelDev *models, unsigned long long *red,
                         unsigned *inv, double *xs, double *ys, double *zs, unsigned *sid, unsigned retired)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ unsigned long long sm[8][256];
    __shared__ int sc_first, sc_last;
    unsigned long long v[8];
    int sc = -1;
    if (i < n) {
        const double4 p = g.pos[i], u = g.vel[i];
        sc = models[g.mdl[i]].scene;
        v[0] = ord64(p.x); v[1] = ord64(p.y); v[2] = ord64(p.z);
        v[3] = ~ord64(p.x); v[4] = ~ord64(p.y); v[5] = ~ord64(p.z);
        v[6] = ord64(len(mk(u.x, u.y, u.z)));
        v[7] = ord64(p.w);
        const unsigned o = g.oid[i];
        const bool dead = (o & OID_DEAD) == OID_DEAD;   // erased by the zone of interest
        inv[dead ? (o & ~OID_DEAD) : scenes[sc].rank_off + o] = i;
        xs[i] = p.x; ys[i] = p.y; zs[i] = p.z;
        sid[i] = (scenes[sc].active < 0 || dead) ? retired : (unsigned)sc;
        if (scenes[sc].active < 0 || dead)   // retired or erased: no part in any grid
            for (int k = 0; k < 8; ++k) v[k] = 0;
    } else {
        for (int k = 0; k < 8; ++k) v[k] = 0;
    }
    const int last_i = min(n, (int)((blockIdx.x + 1) * blockDim.x)) - 1;
    if (threadIdx.x == 0) sc_first = sc;
    if (i == last_i) sc_last = sc;
    __syncthreads();
    if (sc_first != sc_last) {
        // the block straddles scenes (block-uniform branch): every sphere adds its own values
        if (i < n)
            for (int k = 0; k < 8; ++k) atomicMax(&red[(size_t)sc * 8 + k], v[k]);
        return;
    }
    for (int k = 0; k < 8; ++k) sm[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 8; ++k) sm[k][threadIdx.x] = max(sm[k][threadIdx.x], sm[k][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 8 && sc_first >= 0) atomicMax(&red[(size_t)sc_first * 8 + threadIdx.x], sm[threadIdx.x][0]);
}

// Values at given positions of the sorted coordinate arrays (the 2%-trimmed
// span quantiles of every scene): idx[m][2] -> out[m][6].
__global__ void k_quantiles(const double *xs, const double *ys, const double *zs, const long long *idx, int m,
                            double *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const long long lo = idx[2 * t], hi = idx[2 * t + 1];
    out[6 * t + 0] = xs[lo]; out[6 * t + 1] = xs[hi];
    out[6 * t + 2] = ys[lo]; out[6 * t + 3] = ys[hi];
    out[6 * t + 4] = zs[lo]; out[6 * t + 5] = zs[hi];
}

// Exact order statistics of each scene's coordinates by radix selection.
// LeapConsultant::Reset sorts every coordinate to read its 2% and 98%
// elements (LeapConsultant.cpp:139-191); the same elements are found here
// without a sort.  Target t = 6 q + 2 c + side (q: the live scene's index,
// c: x / y / z, side 0 / 1: the lower / upper element) keeps the rank still
// to find and the bits of the element's order-preserving key settled so far.
// Pass p (0..7) counts, per target, the byte at bits 56 - 8p .. 63 - 8p of
// the keys that share the target's settled bits; k_qsel_pick then settles
// that byte.  Only spheres of live scenes count (sid names their slot;
// erased and retiring spheres carry another value).
__device__ inline double unord64(unsigned long long u)
{
    const unsigned long long b = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
    return __longlong_as_double((long long)b);
}

__global__ void k_qsel_hist(Geo g, int n, const unsigned *sid, const int *qmap, int n_slots, const QselTarget *tg,
                            int pass, unsigned *hist)
{
    __shared__ unsigned sh[6][256];
    __shared__ int q_lo, q_hi;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = __lane_id();
    int q = -1;
    if (i < n) {
        const unsigned s = sid[i];
        q = s < (unsigned)n_slots ? qmap[s] : -1;
    }
    for (int k = threadIdx.x; k < 6 * 256; k += blockDim.x) sh[k >> 8][k & 255] = 0u;
    if (threadIdx.x == 0) {
        q_lo = 0x7fffffff;
        q_hi = -1;
    }
    // the block's range of live-scene indices (wavefront reductions, then LDS)
    int lo = q >= 0 ? q : 0x7fffffff, hi = q;
    for (int d = 32; d > 0; d >>= 1) {
        lo = min(lo, __shfl_xor(lo, d, 64));
        hi = max(hi, __shfl_xor(hi, d, 64));
    }
    __syncthreads();
    if (lane == 0) {
        atomicMin(&q_lo, lo);
        atomicMax(&q_hi, hi);
    }
    __syncthreads();
    // one scene in the block (the usual case): counts gather in LDS first
    const bool one = q_lo == q_hi;
    const int shift = 56 - 8 * pass;
    double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
    if (q >= 0) p = g.pos[i];
    const double cc[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const unsigned long long key = ord64(cc[c]);
        const unsigned d = (unsigned)(key >> shift) & 255u;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            bool match = q >= 0;
            if (match && pass > 0) {
                const unsigned long long pre = tg[6 * q + 2 * c + side].prefix;
                match = (key >> (shift + 8)) == (pre >> (shift + 8));
            }
            if (one) {
                // one LDS atomic per distinct byte of the wavefront's matching lanes
                unsigned long long act = __ballot(match);
                while (act) {
                    const int first = __builtin_ctzll(act);
                    const unsigned d0 = __shfl(d, first, 64);
                    const unsigned long long grp = act & __ballot(d == d0);
                    if (lane == first) atomicAdd(&sh[2 * c + side][d0], (unsigned)__popcll(grp));
                    act &= ~grp;
                }
            } else if (match) {
                atomicAdd(&hist[((size_t)6 * q + 2 * c + side) * 256 + d], 1u);
            }
        }
    }
    __syncthreads();
    if (one && q_hi >= 0)
        for (int k = threadIdx.x; k < 6 * 256; k += blockDim.x) {
            const unsigned v = sh[k >> 8][k & 255];
            if (v) atomicAdd(&hist[((size_t)6 * q_hi + (k >> 8)) * 256 + (k & 255)], v);
        }
}

// One wavefront per target: the byte whose counts cover the target's rank,
// then the counters are cleared for the next pass.
__global__ void k_qsel_pick(QselTarget *tg, int m, int pass, unsigned *hist, double *out)
{
    const int t = blockIdx.x * (blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    const int lane = __lane_id();
    if (t >= m) return;   // uniform across the wavefront
    uint4 *hh = reinterpret_cast<uint4 *>(hist + (size_t)t * 256);
    const uint4 b = hh[lane];   // bytes 4 lane .. 4 lane + 3
    const unsigned long long r = tg[t].rank;
    const unsigned long long own = (unsigned long long)b.x + b.y + b.z + b.w;
    unsigned long long incl = own;
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned long long y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
    }
    const unsigned long long over = __ballot(incl > r);
    const int L = over ? __builtin_ctzll(over) : 63;
    unsigned long long cum = __shfl(incl - own, L, 64);
    const unsigned c0 = __shfl(b.x, L, 64), c1 = __shfl(b.y, L, 64), c2 = __shfl(b.z, L, 64);
    int d = 4 * L;
    if (cum + c0 <= r) {
        cum += c0;
        ++d;
        if (cum + c1 <= r) {
            cum += c1;
            ++d;
            if (cum + c2 <= r) {
                cum += c2;
                ++d;
            }
        }
    }
    hh[lane] = make_uint4(0u, 0u, 0u, 0u);
    if (lane == 0) {
        const int shift = 56 - 8 * pass;
        const unsigned long long pre = tg[t].prefix | ((unsigned long long)d << shift);
        tg[t].prefix = pre;
        tg[t].rank = r - cum;
        if (pass == 7) out[t] = unord64(pre);
    }
}

__device__ inline void cell_of(const GridDev &G, d3 p, unsigned &cx, unsigned &cy, unsigned &cz)
{
    cx = (unsigned)((p.x - G.xmin) / G.zr);
    cy = (unsigned)((p.y - G.ymin) / G.zr);
    cz = (unsigned)((p.z - G.zmin) / G.zr);
}
__device__ inline long long lin_of(const GridDev &G, long long cx, long long cy, long long cz)
{
    return ((cz + 1) * G.ny + (cy + 1)) * G.nx + (cx + 1);
}

// Memory order of the spheres: the Hilbert curve through the cells, so that
// the cells around a sphere sit close in memory (consecutive cells of the
// curve are always face neighbours, where the Morton order jumps at every
// power-of-two boundary); each sphere's list order is independent of it
// (k_outer orders by the comparator key).  Keys are produced in DO-id order
// and the sort is stable, so each cell's spheres stay in id order as the
// reference's per-cell vectors do (LeapConsultant.h:170-188).
__device__ inline unsigned long long spread3(unsigned v)
{
    unsigned long long x = v & 0x1ffffull;   // 17 bits per axis (checked on the host)
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}
// Hilbert index of cell (x, y, z), 17 bits per axis: J. Skilling's transform
// of the axes to the "transposed" index ("Programming the Hilbert curve",
// AIP Conf. Proc. 707, 381 (2004)), whose bits interleave to the index.
__device__ inline unsigned long long hilbert3(unsigned x, unsigned y, unsigned z)
{
    unsigned X[3] = {x, y, z};
    constexpr unsigned M = 1u << 16;
    for (unsigned Q = M; Q > 1; Q >>= 1) {   // inverse undo
        const unsigned P = Q - 1;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const unsigned t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];   // Gray encode
    X[2] ^= X[1];
    unsigned t = 0;
    for (unsigned Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    X[0] ^= t;
    X[1] ^= t;
    X[2] ^= t;
    return (spread3(X[0]) << 2) | (spread3(X[1]) << 1) | spread3(X[2]);
}

__global__ void k_cellkey(Geo g, int n_ranks, const SceneDev *scenes, const ModelDev *models, const unsigned *inv,
                          unsigned long long retired_key, int hbits, unsigned long long *keys, unsigned *vals)
{
    // key = scene slot << hbits | Hilbert index of the cell (below 2^hbits for
    // every placed scene's grid); ranks run scene by scene in DO-id order,
    // retired scenes last
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_ranks) return;
    const unsigned i = inv[t];
    const int sc = models[g.mdl[i]].scene;
    const SceneDev &S = scenes[sc];
    vals[t] = i;
    if (S.active < 0 || (g.oid[i] & OID_DEAD) == OID_DEAD) {
        keys[t] = retired_key;
        return;
    }
    const double4 p = g.pos[i];
    unsigned cx, cy, cz;
    cell_of(S.G, mk(p.x, p.y, p.z), cx, cy, cz);
    keys[t] = ((unsigned long long)sc << hbits) | hilbert3(cx, cy, cz);
}

// Gather state into cell order; per-sphere safe distance
// SafetyFactor*(2*range + dt_up*|v| + 0.5*dt_up^2*|g|) (LeapConsultant.h:117-119).
__global__ void k_reorder(Geo src, Geo dst, int n, int ld, const unsigned *perm, const SceneDev *scenes,
                          const ModelDev *models, double *safe, int *lin, const double *imp_src, double *imp_dst)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const unsigned p = perm[s];
    const double4 x = src.pos[p];
    double4 v = src.vel[p];
    const int m = src.mdl[p];
    const SceneDev &S = scenes[models[m].scene];
    const GridDev &G = S.G;
    dst.pos[s] = x;
    safe[s] = G.sf * (2.0 * x.w + G.safe_v * len(mk(v.x, v.y, v.z)) + G.safe_g);
    v.w = 0.0;
    dst.vel[s] = v;
    dst.avl[s] = src.avl[p];
    for (int c = 0; c < 6; ++c) dst.ori[(size_t)c * ld + s] = src.ori[(size_t)c * ld + p];
    dst.oid[s] = src.oid[p];
    dst.mdl[s] = m;
    if (imp_src)   // last-step impulses follow their sphere
        for (int c = 0; c < 6; ++c) imp_dst[(size_t)c * ld + s] = imp_src[(size_t)c * ld + p];
    unsigned cx, cy, cz;
    cell_of(G, mk(x.x, x.y, x.z), cx, cy, cz);
    lin[s] = (int)(S.cell_off + lin_of(G, cx, cy, cz));
}

__global__ void k_fill_int(int *p, long long n, int v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_cellstart(const int *lin, int n, int *cstart, int *cend)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    if (s == 0 || lin[s] != lin[s - 1]) cstart[lin[s]] = s;
    if (s == n - 1 || lin[s] != lin[s + 1]) cend[lin[s]] = s + 1;
}

__device__ inline bool accept_wall(const OuterArgs &a, const SceneDev &S, d3 xs, double safe_s, int gs, int w)
{
    const WallDev &W = a.walls[w];
    if (a.t.laws[S.law_off + gs * S.n_groups + W.group].law == LAW_NONE) return false;
    const GridDev &G = S.G;
    const double sw = G.sf * (2.0 * W.range + G.safe_v * len(W.vel) + G.safe_g);
    return !(domap_wall_distance(xs, W) > safe_s + sw);
}

// Candidate partners of sphere s in the reference's IactPairTab order: the
// groups "current cell = c" of the Leap traversal are visited in comparator
// order; within the own cell the 13 half-shell cells, the self cell and the
// global DOs follow in reference order (LeapConsultant.h:207-444).  Each
// accepted pair also carries the marks of its contact depth for the
// neighbour list.
// The scene's grid and law table are read once into registers and the
// arrays are declared unaliased, so each candidate test costs one gather of
// the partner's position and safe distance, issued one candidate ahead.
template <bool HOLE>
__global__ __launch_bounds__(128) void k_outer(OuterArgs a)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.n) return;
    const double4 *__restrict__ pos = a.g.pos;
    const double *__restrict__ safe = a.safe;
    const int *__restrict__ mdl = a.g.mdl;
    const ModelDev *__restrict__ models = a.t.models;
    const LawDev *__restrict__ laws = a.t.laws;
    unsigned *__restrict__ nbr = a.nbr;
    const size_t ld = (size_t)a.ld;
    const unsigned ko = (unsigned)a.ko;
    const bool single_group = a.single_group != 0;

    const double4 p = pos[s];
    const d3 xs = mk(p.x, p.y, p.z);
    const double safe_s = safe[s];
    const ModelDev &ms = models[mdl[s]];
    const int gs = ms.group, scene = ms.scene;
    const SceneDev &Sg = a.scenes[scene];
    const GridDev G = Sg.G;
    const int law_off = Sg.law_off, n_groups = Sg.n_groups, w_begin = Sg.w_begin, w_end = Sg.w_end;
    const int *__restrict__ cstart = a.cstart + Sg.cell_off;
    const int *__restrict__ cend = a.cend + Sg.cell_off;
    // partner groups with a sphere-sphere law (DOWorld::GetIactModel != 0)
    const bool mask_ok = n_groups <= 64;
    unsigned long long lawmask = 0;
    if (mask_ok)
        for (int g = 0; g < n_groups; ++g)
            if (laws[law_off + gs * n_groups + g].law != LAW_NONE) lawmask |= 1ull << g;
    const bool any_sphere_law = !single_group || !mask_ok || ((lawmask >> gs) & 1ull);

    unsigned cx, cy, cz;
    cell_of(G, xs, cx, cy, cz);
    const double skin = Sg.skin;
    unsigned cnt = 0, near_cnt = 0;   // candidates; those within the skin (the neighbour list's length)
    auto emit = [&](unsigned j) {
        if (cnt < ko) nbr[(size_t)cnt * ld + s] = j;
        ++cnt;
        near_cnt += (j & NBR_NEAR) ? 1u : 0u;
    };
    // the contact depth of an accepted pair (CDSphere_Sphere's, the same bits
    // from either side) marks it for the neighbour list: within the skin of
    // contact, in contact
    auto marks = [&](double depth) -> unsigned {
        return (depth >= -skin ? NBR_NEAR : 0u) | (depth > 0.0 ? NBR_ACTIVE : 0u);
    };
    auto sphere_marks = [&](double4 pj) -> unsigned {
        const d3 xj = mk(pj.x, pj.y, pj.z);
        const d3 dx = xj - xs;
        const double lim = p.w + pj.w + skin;
        if (dot(dx, dx) > lim * lim * (1.0 + 1e-9)) return 0u;   // surely beyond the skin
        return marks(p.w + pj.w - len(dx));
    };
    // DOMap::CalDistance of the pair against safe_i + safe_j: the squared
    // distance decides every pair away from the threshold; near it the
    // reference's own test `len(d) > safe_i + safe_j` decides
    auto near = [&](double4 pj, double sj) -> bool {
        const d3 d = xs - mk(pj.x, pj.y, pj.z);
        const double lim = safe_s + sj;
        const double d2 = dot(d, d), l2 = lim * lim;
        if (d2 > l2 * (1.0 + 1e-9)) return false;
        if (d2 < l2 * (1.0 - 1e-9)) return true;
        return !(len(d) > lim);
    };
    auto has_law = [&](int j) -> bool {
        if (single_group) return true;   // any_sphere_law holds
        const int gj = models[mdl[j]].group;
        return mask_ok ? ((lawmask >> gj) & 1ull) != 0 : laws[law_off + gs * n_groups + gj].law != LAW_NONE;
    };
    // a cell's spheres in groups of 4: the gathers of a group are in flight together
    auto scan_cell = [&](long long x, long long y, long long z, bool self) {
        if (!any_sphere_law) return;
        if (x < -1 || y < -1 || z < -1 || x > G.nx - 3 || y > G.ny - 3 || z > G.nz - 3) return;
        const long long l = lin_of(G, x, y, z);
        const int b = cstart[l], e = cend[l];
        for (int j0 = b; j0 < e; j0 += 4) {
            double4 pj[4];
            double sj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int jj = j0 + u < e ? j0 + u : e - 1;
                pj[u] = pos[jj];
                sj[u] = safe[jj];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + u;
                if (j >= e) break;
                if (self && j == s) continue;
                if (has_law(j) && near(pj[u], sj[u])) emit((unsigned)j | sphere_marks(pj[u]));
            }
        }
    };
    for (int q = 0; q < 14; ++q) {
        const int slot = (int)((G.order >> (4 * q)) & 15ull);
        if (slot == 13) {
            for (int k = 0; k < 13; ++k)
                scan_cell((long long)cx + NB13[k][0], (long long)cy + NB13[k][1], (long long)cz + NB13[k][2], false);
            scan_cell(cx, cy, cz, true);
            for (int w = w_begin; w < w_end; ++w)
                if (accept_wall(a, Sg, xs, safe_s, gs, w)) {
                    Contact c;
                    cd_wall_at<HOLE>(xs, p.w, a.walls[w], a.holes, w, c);
                    emit((unsigned)(a.ld + w) | marks(c.depth));
                }
        } else {
            // the cell whose half shell holds s's cell, if it exists (LeapConsultant.h:207-444)
            const long long ox = (long long)cx - NB13[slot][0], oy = (long long)cy - NB13[slot][1],
                            oz = (long long)cz - NB13[slot][2];
            if (ox < 0 || oy < 0 || oz < 0 || ox >= G.nx - 2 || oy >= G.ny - 2 || oz >= G.nz - 2) continue;
            scan_cell(ox, oy, oz, false);
        }
    }
    a.count[s] = cnt;
    if (cnt > ko) {
        atomicMax(&a.flags[5], (int)cnt);
        if (a.ovf) atomicOr(&a.ovf[scene], 1);
    }
    if (near_cnt > (unsigned)a.ki) atomicMax(&a.flags[6], (int)near_cnt);
}
template __global__ void k_outer<false>(OuterArgs);
template __global__ void k_outer<true>(OuterArgs);

// ---------------------------------------------------------------------------
// Candidate list by cell pairs: periodic boundaries, NBS
// ---------------------------------------------------------------------------
// The Leap traversal with wrapped half-shell cells (LeapConsultant.h:146-444)
// visits cell pairs that no fixed per-sphere order describes (a cell on a
// border can be paired with many cells of the opposite side), so the list is
// built as the reference builds it: the occupied cells in comparator order,
// each emitting its pairs in reference order (13 lookups, its own pairs, the
// global DOs; cell i-major, partner j-minor), then every pair goes to both
// spheres' lists by a stable sort on the sphere.  A pair the traversal
// reaches twice is listed twice, as the reference's IactPairTab lists it.
// The NBSConsultant's traversal (NBSConsultant.cpp:85-316) is built the same
// way: zones in (x, y, z) order within the zone counts (a sphere on the
// upper bound of an axis sits one zone beyond and is reached only as a
// neighbour), 12 lookups, its own pairs, the global DOs, then the 13th
// lookup (+1, 0, -1); its distances take no periodic image
// (NBSConsultant.cpp:330).

// DOMap::CalDistance's distance vector for sphere-plate / sphere-cylinder
// (DOMap.cpp:320-414), before the boundary's minimum image and the length.
__device__ inline d3 domap_wall_vec(d3 ca, const WallDev &w)
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
        return caps - ca;
    }
    double hh = 0.5 * w.a1;
    double d = dot(ca - w.pos, w.oz);
    d3 cap = w.pos + muls(w.oz, d);
    if (d < hh && d > -hh) return cap - ca;
    if (d >= hh) return (w.pos + muls(w.oz, hh)) - ca;
    return (w.pos - muls(w.oz, hh)) - ca;
}

// cell coordinates of the comparator key (each coordinate + 1, 21 bits)
__device__ inline void pbc_key_cell(const GridDev &G, unsigned long long key, unsigned c[3])
{
    c[G.perm[0]] = (unsigned)(key >> 42) - 1u;
    c[G.perm[1]] = (unsigned)((key >> 21) & 0x1fffffull) - 1u;
    c[G.perm[2]] = (unsigned)(key & 0x1fffffull) - 1u;
}

// occupied cells: the first sphere of each cell run of the sorted layout
__global__ void k_pbc_cells(Geo g, int n, const int *lin, const SceneDev *scene, unsigned long long *ckey,
                            unsigned *cfirst, unsigned *ncell)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    if (s > 0 && lin[s] == lin[s - 1]) return;
    const GridDev &G = scene->G;
    const double4 p = g.pos[s];
    unsigned c[3];
    cell_of(G, mk(p.x, p.y, p.z), c[0], c[1], c[2]);
    const unsigned q = atomicAdd(ncell, 1u);
    ckey[q] = ((unsigned long long)(c[G.perm[0]] + 1u) << 42) | ((unsigned long long)(c[G.perm[1]] + 1u) << 21) |
              (unsigned long long)(c[G.perm[2]] + 1u);
    cfirst[q] = (unsigned)s;
}

// Pairs of sorted cell r in reference order; counts them, or writes them
// to out (EMIT): (sphere i, partner j | marks), j = ld + w for a wall slot.
template <bool EMIT>
__device__ unsigned pbc_cell_pairs(const PbcArgs &a, int r, uint2 *out)
{
    const SceneDev &S = *a.scene;
    const GridDev &G = S.G;
    const PbcDev *pb = a.pbc;                     // contact depths (the marks)
    const PbcDev *pbd = a.nbs ? nullptr : a.pbc;  // candidate distances and wrapped lookups
    const double4 *__restrict__ pos = a.g.pos;
    const double *__restrict__ safe = a.safe;
    const int *__restrict__ mdl = a.g.mdl;
    const ModelDev *__restrict__ models = a.t.models;
    const LawDev *__restrict__ laws = a.t.laws;
    const int *cstart = a.cstart + S.cell_off, *cend = a.cend + S.cell_off;
    unsigned c[3];
    pbc_key_cell(G, a.ckey[r], c);
    if (a.nbs && (c[0] >= a.nc[0] || c[1] >= a.nc[1] || c[2] >= a.nc[2])) return 0u;
    const int b = (int)a.cfirst[r];
    const int e = cend[lin_of(G, c[0], c[1], c[2])];
    const double skin = S.skin;
    unsigned cnt = 0;
    auto cell_range = [&](long long x, long long y, long long z, int &tb, int &te) -> bool {
        if (x < -1 || y < -1 || z < -1 || x > G.nx - 3 || y > G.ny - 3 || z > G.nz - 3) return false;
        const long long l = lin_of(G, x, y, z);
        tb = cstart[l];
        te = cend[l];
        return te > tb;
    };
    auto law_of = [&](int gi, int gj) { return laws[S.law_off + gi * S.n_groups + gj].law; };
    // BuildIactTab's test for spheres i, j (LeapConsultant.cpp:197-262)
    auto pair_ss = [&](int i, int j) {
        if (i == j) return;   // Assembler::CreateInteraction never pairs a DO with itself
        if (law_of(models[mdl[i]].group, models[mdl[j]].group) == LAW_NONE) return;
        const double4 pi = pos[i], pj = pos[j];
        const d3 xi = mk(pi.x, pi.y, pi.z), xj = mk(pj.x, pj.y, pj.z);
        if (len(bc_diff(pbd, xi - xj)) > safe[i] + safe[j]) return;
        if (EMIT) {
            const double depth = pi.w + pj.w - len(bc_diff(pb, xj - xi));
            const unsigned m = (depth >= -skin ? NBR_NEAR : 0u) | (depth > 0.0 ? NBR_ACTIVE : 0u);
            out[cnt] = make_uint2((unsigned)i, (unsigned)j | m);
        }
        ++cnt;
    };
    auto lookup = [&](int k) {
        int tb = 0, te = 0;
        bool found = cell_range((long long)c[0] + NB13[k][0], (long long)c[1] + NB13[k][1],
                                (long long)c[2] + NB13[k][2], tb, te);
        if (!found && pbd && (pbd->pax[0] || pbd->pax[1] || pbd->pax[2])) {
            // the wrapped half-shell cell (LeapConsultant.h:199-444)
            bool border = false, dup = false;
            unsigned w[3];
            for (int d = 0; d < 3; ++d) {
                const int o = NB13[k][d];
                if (o < 0) {
                    border = border || c[d] == 0u;
                    w[d] = pbd->pax[d] ? pbd->maxid[d] : c[d] - 1u;
                    dup = dup || w[d] == c[d] + 1u;
                } else if (o > 0) {
                    border = border || c[d] == pbd->maxid[d];
                    w[d] = pbd->pax[d] ? 0u : c[d] + 1u;
                    dup = dup || w[d] == c[d] - 1u;
                } else {
                    w[d] = c[d];
                }
            }
            if (border && !dup) found = cell_range(w[0], w[1], w[2], tb, te);
        }
        if (found)
            for (int i = b; i < e; ++i)
                for (int j = tb; j < te; ++j) pair_ss(i, j);
    };
    for (int k = 0; k < (a.nbs ? 12 : 13); ++k) lookup(k);
    for (int i = b; i < e; ++i)
        for (int j = i + 1; j < e; ++j) pair_ss(i, j);
    for (int i = b; i < e; ++i) {
        const int gi = models[mdl[i]].group;
        const double4 pi = pos[i];
        const d3 xi = mk(pi.x, pi.y, pi.z);
        for (int w = S.w_begin; w < S.w_end; ++w) {
            const WallDev &W = a.walls[w];
            if (law_of(gi, W.group) == LAW_NONE) continue;
            const double sw = G.sf * (2.0 * W.range + G.safe_v * len(W.vel) + G.safe_g);
            if (len(bc_diff(pbd, domap_wall_vec(xi, W))) > safe[i] + sw) continue;
            if (EMIT) {
                Contact cw;
                cd_wall_at<true>(xi, pi.w, W, a.holes, w, cw);
                const unsigned m = (cw.depth >= -skin ? NBR_NEAR : 0u) | (cw.depth > 0.0 ? NBR_ACTIVE : 0u);
                out[cnt] = make_uint2((unsigned)i, (unsigned)(a.ld + w) | m);
            }
            ++cnt;
        }
    }
    if (a.nbs) lookup(12);   // NBSConsultant.cpp:268-271
    return cnt;
}

__global__ void k_pbc_count(PbcArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.ncell) return;
    a.pcount[r] = pbc_cell_pairs<false>(a, r, nullptr);
}

__global__ void k_pbc_emit(PbcArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.ncell) return;
    pbc_cell_pairs<true>(a, r, a.pairs + a.poff[r]);
}

// two list entries per sphere pair (one for a wall pair), keyed by the
// sphere whose list takes them; the value is the pair and its side
__global__ void k_pbc_entries(const uint2 *pairs, unsigned np, int ld, unsigned *key, unsigned *val)
{
    const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const uint2 q = pairs[p];
    const unsigned j = q.y & PARTNER_MASK;
    key[2 * p] = q.x;
    val[2 * p] = 2 * p;
    key[2 * p + 1] = j < (unsigned)ld ? j : 0xffffffffu;
    val[2 * p + 1] = 2 * p + 1;
}

__global__ void k_pbc_segstart(const unsigned *key, unsigned ne, unsigned *first)
{
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const unsigned s = key[e];
    if (s == 0xffffffffu) return;
    if (e == 0 || key[e - 1] != s) first[s] = e;
}

// each sphere's entries, in pair order, into its candidate list
__global__ void k_pbc_lists(const unsigned *key, const unsigned *val, unsigned ne, const uint2 *pairs,
                            const unsigned *first, int ld, int ko, int ki, unsigned *count, unsigned *nbr,
                            int *flags)
{
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const unsigned s = key[e];
    if (s == 0xffffffffu) return;
    const unsigned t = e - first[s];
    const unsigned v = val[e];
    const uint2 q = pairs[v >> 1];
    const unsigned ent = (v & 1u) ? (q.x | (q.y & ~PARTNER_MASK)) : q.y;
    if (t < (unsigned)ko) nbr[(size_t)t * ld + s] = ent;
    if (e + 1 == ne || key[e + 1] != s) {
        count[s] = t + 1;
        if ((int)(t + 1) > ko) atomicMax(&flags[5], (int)(t + 1));
        if ((int)(t + 1) > ki) {
            // the sphere's entries within the skin (its neighbour list's length),
            // from the pairs this kernel reads (the entries' marks are the pair's)
            int near_cnt = 0;
            for (unsigned e2 = first[s]; e2 <= e; ++e2) near_cnt += (pairs[val[e2] >> 1].y & NBR_NEAR) ? 1 : 0;
            if (near_cnt > ki) atomicMax(&flags[6], near_cnt);
        }
    }
}

// Consistency of the device lists: every index the step, wall and rebuild
// kernels form from them lies where it must (checked before it is used):
// each sphere's model and scene, its neighbour-list count and entries
// (a placed sphere of the same scene, or a wall slot of the scene's range),
// its candidate-list count and entries likewise, its constrained-wall slots
// and pair slots.  out[kind] counts violations; out[8..10] hold the first
// one (kind, sphere, entry).
__global__ void k_check_lists(Geo g, Ell ell, int n, int ld, int ki, int ko, const unsigned *o_cnt,
                              const unsigned *o_nbr, const ModelDev *models, int n_models, const SceneDev *scenes,
                              int scap, const WallDev *walls, int wcap, const int *cw_w, const int *cw_p,
                              long long cw_cap, unsigned long long *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    auto bad = [&](int kind, int k) {
        if (atomicAdd(&out[kind], 1ull) == 0ull && atomicCAS(&out[8], 0ull, (unsigned long long)kind + 1) == 0ull) {
            out[9] = (unsigned long long)i;
            out[10] = (unsigned long long)(long long)k;
        }
    };
    const int m = g.mdl[i];
    if (m < 0 || m >= n_models) { bad(0, -1); return; }
    const int sc = models[m].scene;
    if (sc < 0 || sc >= scap) { bad(1, -1); return; }
    const SceneDev &S = scenes[sc];
    if (S.w_begin < 0 || S.w_end > wcap || S.w_begin > S.w_end) { bad(1, -2); return; }
    auto partner_ok = [&](unsigned j) -> bool {
        if (j < (unsigned)ld) {
            if (j >= (unsigned)n) return false;
            const int mj = g.mdl[j];
            return mj >= 0 && mj < n_models && models[mj].scene == sc;
        }
        const unsigned w = j - (unsigned)ld;
        return w >= (unsigned)S.w_begin && w < (unsigned)S.w_end && walls[w].scene == sc;
    };
    const unsigned hdr = ell.hdr[i];
    const int h = hdr >> 7, cnt = hdr & 127;
    if (cnt > ki) bad(2, cnt);
    Bits<2> used;
    for (int k = 0; k < min(cnt, ki); ++k) {
        const size_t x = ((size_t)(h * ki + k)) * (size_t)ld + i;
        if (!partner_ok(ell.ent[x] & PARTNER_MASK)) bad(3, k);
        if (ell.ent[x] & F_CONTACT) {
            // a contact's history slot: within the capacity, no two contacts share one
            const unsigned sl = ell.slot[x];
            if (sl >= (unsigned)ki || sl >= (unsigned)KI_MAX || used.has((int)sl)) bad(4, k);
            else used.set((int)sl);
        }
    }
    const unsigned oc = o_cnt[i];
    if (oc > (unsigned)ko) bad(5, (int)oc);
    for (unsigned t = 0; t < min(oc, (unsigned)ko); ++t)
        if (!partner_ok(o_nbr[(size_t)t * ld + i] & PARTNER_MASK)) bad(6, (int)t);
    if (cw_w)
        for (int q = 0; q < CW_SLOTS; ++q) {
            const int w = cw_w[(size_t)q * ld + i];
            if (w == -1) continue;
            if (w < S.w_begin || w >= S.w_end || !walls[w].constrained) bad(7, q);
            const int pq = cw_p[(size_t)q * ld + i];
            if (pq < 0 || (long long)pq >= cw_cap) bad(7, CW_SLOTS + q);
        }
}

// Consultant::GetIactNum: sphere pairs are listed from both sides, walls once.
__global__ void k_count_iact(const unsigned *cnt, const unsigned *nbr, int b, int e, int ld, int ko,
                             unsigned long long *out)
{
    const int s = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= e) return;
    unsigned sph = 0, wal = 0;
    for (unsigned t = 0; t < cnt[s] && t < (unsigned)ko; ++t) {
        if ((nbr[(size_t)t * ld + s] & PARTNER_MASK) < (unsigned)ld) ++sph;
        else ++wal;
    }
    atomicAdd(&out[0], (unsigned long long)sph);
    atomicAdd(&out[1], (unsigned long long)wal);
}

// Neighbour lists at a rebuild, with IactRecordTab's history carry: a pair
// keeps its ImpactStatus if it was listed and is active now
// (RebuildIactRecordTab + Initiate, NBSConsultant.cpp:376-395,
// SimMediator.cpp:690-707); everything else starts from ImpactStatus().
__global__ __launch_bounds__(128) void k_inner_build(InnerArgs a)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    extern __shared__ unsigned old_oid[];   // [ki_old][blockDim]
    if (s >= a.n) return;
    const unsigned *__restrict__ oid = a.g.oid;
    const unsigned *__restrict__ nbr = a.nbr;
    const unsigned *__restrict__ ent_old = a.eold.ent;
    unsigned *__restrict__ ent_new = a.ell.ent;
    const size_t ld = (size_t)a.ld;
    const unsigned p = a.perm[s];
    const unsigned hdr = a.eold.hdr[p];
    const int h_old = hdr >> 7, c_old = hdr & 127;
    int n_hist = 0;   // the new list's contacts take history slots 0, 1, ... in list order
    for (int k = 0; k < c_old; ++k) {
        const unsigned j = ent_old[eidx(h_old, k, a.ki_old, a.ld, p)] & PARTNER_MASK;
        old_oid[k * blockDim.x + threadIdx.x] = j < (unsigned)a.ld ? a.gold.oid[j] : 0x80000000u + (j - a.ld);
    }
    const unsigned my_oid = oid[s];
    const unsigned cnt = a.cnt[s];
    int q = 0;
    // the candidates the rebuild marked as within the skin, in list order;
    // their ranks are gathered 4 at a time
    for (unsigned t0 = 0; t0 < cnt; t0 += 4) {
        unsigned jv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) jv[u] = t0 + u < cnt ? nbr[(size_t)(t0 + u) * ld + s] : 0u;
        unsigned kj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const unsigned j = jv[u] & PARTNER_MASK;
            kj[u] = !(jv[u] & NBR_NEAR) ? 0u : (j < (unsigned)a.ld ? oid[j] : 0x80000000u + (j - a.ld));
        }
        bool full = false;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (!(jv[u] & NBR_NEAR)) continue;
            const unsigned j = jv[u] & PARTNER_MASK;
            const unsigned key = kj[u];
            const bool slave = j >= (unsigned)a.ld || my_oid < key;
            if (q >= a.ki) {
                atomicOr(&a.flags[3], ERR_INNER_CAP);
                full = true;
                break;
            }
            unsigned f = F_DEFAULT;
            Hist hs;
            hs.s = mk(0, 0, 0);
            hs.aux = 0.0;
            if (jv[u] & NBR_ACTIVE) {
                for (int k = 0; k < c_old; ++k)
                    if (old_oid[k * blockDim.x + threadIdx.x] == key) {
                        f = (ent_old[eidx(h_old, k, a.ki_old, a.ld, p)] & ~PARTNER_MASK & ~F_SLAVE) | F_NOBACKUP;
                        if (f & F_CONTACT) load_hist(a.eold, a.eold.slot[eidx(h_old, k, a.ki_old, a.ld, p)], a.ld, p, hs);
                        break;
                    }
            }
            if (slave) f |= F_SLAVE;
            ent_new[eidx(0, q, a.ki, a.ld, s)] = j | f;
            if (f & F_CONTACT) {
                a.ell.slot[eidx(0, q, a.ki, a.ld, s)] = (unsigned char)n_hist;
                store_hist(a.ell, n_hist, a.ld, s, hs);
                ++n_hist;
            }
            ++q;
        }
        if (full) break;
    }
    a.ell.hdr[s] = (unsigned char)q;
    atomicMax(&a.flags[4], q);
}

// ---------------------------------------------------------------------------
// Constrained walls
// ---------------------------------------------------------------------------
// A constrained wall's impulse is the sum of its pairs' master impulses in
// IactPairTab order (IactContainer::CalculateImpact, IactContainer.cpp:33-37,
// adding with DiscreteObject::AddImpact), then the field and external
// impulses (DOContainer.cpp:20-50).  Its pairs in that order are the spheres
// that list it as a candidate, ordered by the comparator key of their cell at
// the rebuild and, within a cell, by their in-cell (DO-id) order, which is
// the memory order: k_cw_count / k_cw_emit and a stable sort lay them down.

// Comparator key of a sphere's cell, 17 bits per axis in priority order.
__device__ inline unsigned long long cmp_key17(const GridDev &G, d3 x)
{
    unsigned c[3];
    cell_of(G, x, c[0], c[1], c[2]);
    return ((unsigned long long)c[G.perm[0]] << 34) | ((unsigned long long)c[G.perm[1]] << 17) |
           (unsigned long long)c[G.perm[2]];
}

__global__ void k_cw_count(int n, int ld, int ko, const unsigned *cnt, const unsigned *nbr, const WallDev *walls,
                           unsigned *out, int *flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        out[n] = 0;
        return;
    }
    const unsigned e = cnt[i] < (unsigned)ko ? cnt[i] : (unsigned)ko;
    unsigned c = 0;
    for (unsigned t = 0; t < e; ++t) {
        const unsigned j = nbr[(size_t)t * ld + i] & PARTNER_MASK;
        if (j >= (unsigned)ld && walls[j - ld].constrained) ++c;
    }
    if (c > (unsigned)CW_SLOTS) {   // a sphere near more constrained walls than it has pair slots
        atomicOr(&flags[3], ERR_CW);
        c = CW_SLOTS;
    }
    out[i] = c;
}

__global__ void k_cw_emit(int n, int ld, int ko, const unsigned *cnt, const unsigned *nbr, const WallDev *walls,
                          Geo g, const ModelDev *models, const SceneDev *scenes, const unsigned *off,
                          unsigned long long *key, unsigned *val)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned b = off[i], lim = off[i + 1];
    if (b == lim) return;
    const double4 p = g.pos[i];
    const unsigned long long ck = cmp_key17(scenes[models[g.mdl[i]].scene].G, mk(p.x, p.y, p.z));
    const unsigned e = cnt[i] < (unsigned)ko ? cnt[i] : (unsigned)ko;
    unsigned q = 0;
    for (unsigned t = 0; t < e && b + q < lim; ++t) {
        const unsigned j = nbr[(size_t)t * ld + i] & PARTNER_MASK;
        if (j >= (unsigned)ld && walls[j - ld].constrained) {
            key[b + q] = ((unsigned long long)(j - ld) << 51) | ck;
            val[b + q] = ((unsigned)i << 2) | q;
            ++q;
        }
    }
}

__global__ void k_cw_finalize(long long np, const unsigned long long *key, const unsigned *val, int ld, int *cw_w,
                              int *cw_p, unsigned *stamp, int *beg, int *end)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const unsigned v = val[p];
    const int i = (int)(v >> 2), q = (int)(v & 3u);
    const int w = (int)(key[p] >> 51);
    cw_w[(size_t)q * ld + i] = w;
    cw_p[(size_t)q * ld + i] = (int)p;
    stamp[p] = 0xffffffffu;
    if (p == 0 || (int)(key[p - 1] >> 51) != w) beg[w] = (int)p;
    if (p == np - 1 || (int)(key[p + 1] >> 51) != w) end[w] = (int)p + 1;
}

// Each constrained wall's impulse of the step: its pairs' master impulses
// (those solved in this step carry its stamp) in pair order, then the field
// impulse g dt m* and the external impulse dt F (DOContainer.cpp:20-50).
__global__ void k_cw_sum(const WallDev *walls, int nw, const SceneDev *scenes, const int *beg, const int *end,
                         const double *val, const unsigned *stamp, unsigned cur, double *imp)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const WallDev &W = walls[w];
    if (W.scene < 0 || !W.constrained) return;
    const SceneDev &S = scenes[W.scene];
    if (S.active <= 0) return;
    d3 I = mk(0, 0, 0);
    for (int p = beg[w]; p < end[w]; ++p)
        if (stamp[p] == cur) I = I + mk(val[3 * (size_t)p], val[3 * (size_t)p + 1], val[3 * (size_t)p + 2]);
    const d3 gdt = muls(mk(S.gx, S.gy, S.gz), S.dt);
    I = I + muls(gdt, W.msudo);
    I = I + smul(S.dt, mk(W.fext[0], W.fext[1], W.fext[2]));
    imp[3 * (size_t)w] = I.x;
    imp[3 * (size_t)w + 1] = I.y;
    imp[3 * (size_t)w + 2] = I.z;
}

// IactRecordTab restore (SimMediator::Initiate, SimMediator.cpp:690-707, via
// Consultant::RetrieveImpactStatus): every listed pair whose (lo, hi) DO ids
// have a record takes the record's status, as a rebuild's history carry does
// (flags, shear force, Kn / initial velocity; the pending Backup is skipped).
// Records are sorted by key lo << 32 | hi; of equal keys the last applies.
struct RestoreRec {
    unsigned flags, pad;   // F_CONTACT | F_LASTCONTACT | F_BOND | F_LASTBOND
    double kn, init_vel, shear[3];
};

__global__ void k_restore_history(Geo g, Ell ell, int b, int e, int ld, int ki, const unsigned *sphere_id,
                                  const unsigned *wall_id, const unsigned long long *keys, const RestoreRec *recs,
                                  long long nrec, const ModelDev *models, const SceneDev *scenes, const LawDev *laws,
                                  const WallDev *walls)
{
    const int i = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e) return;
    const unsigned hdr = ell.hdr[i];
    const int h = hdr >> 7, cnt = hdr & 127;
    const ModelDev &mi = models[g.mdl[i]];
    const SceneDev &S = scenes[mi.scene];
    const unsigned id = sphere_id[g.oid[i]];
    // history slots of the listed contacts; a restored contact keeps its
    // entry's slot or takes a free one
    Bits<2> used;
    for (int k = 0; k < cnt; ++k) {
        const size_t x = eidx(h, k, ki, ld, i);
        if (ell.ent[x] & F_CONTACT) used.set(ell.slot[x]);
    }
    for (int k = 0; k < cnt; ++k) {
        const size_t x = eidx(h, k, ki, ld, i);
        const unsigned ent = ell.ent[x];
        const unsigned j = ent & PARTNER_MASK;
        unsigned pid;
        int gj;
        if (j < (unsigned)ld) {
            pid = sphere_id[g.oid[j]];
            gj = models[g.mdl[j]].group;
        } else {
            pid = wall_id[j - ld];
            gj = walls[j - ld].group;
        }
        const unsigned lo = id < pid ? id : pid, hi = id < pid ? pid : id;
        const unsigned long long key = ((unsigned long long)lo << 32) | hi;
        long long l = 0, r = nrec;   // first record with key > key; the last of equal keys wins
        while (l < r) {
            const long long m = (l + r) >> 1;
            if (keys[m] <= key) l = m + 1;
            else r = m;
        }
        if (l == 0 || keys[l - 1] != key) continue;
        const RestoreRec &R = recs[l - 1];
        const unsigned f = (R.flags & F_STATUS) | (ent & F_SLAVE) | F_NOBACKUP;
        ell.ent[x] = j | f;
        const int law = laws[S.law_off + mi.group * S.n_groups + gj].law;
        if (!(f & F_CONTACT)) continue;   // history words are read only for contacts
        Hist hs;
        hs.s = mk(R.shear[0], R.shear[1], R.shear[2]);
        hs.aux = law == LAW_H ? R.kn : (law == LAW_HERTZ ? R.init_vel : 0.0);
        const int sl = (ent & F_CONTACT) ? (int)ell.slot[x] : take_slot(used);
        ell.slot[x] = (unsigned char)sl;
        store_hist(ell, sl, ld, i, hs);
    }
}

// ---------------------------------------------------------------------------
// Zone of interest (Consultant::CleanUp, Consultant.cpp:264-319)
// ---------------------------------------------------------------------------

// Spheres of stepping scenes outside their zone of interest (Boundary::
// InBoundary, Boundary.cpp:161-191): flagged, counted per scene.
__global__ void k_zoi_mark(Geo g, int n, const SceneDev *scenes, const ModelDev *models, unsigned char *dead,
                           int *count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int sc = models[g.mdl[i]].scene;
    const SceneDev &S = scenes[sc];
    unsigned char d = 0;
    if (S.zon && S.active > 0) {
        const double4 p = g.pos[i];
        const double c[3] = {p.x, p.y, p.z};
        for (int k = 0; k < 3; ++k)
            if (((S.zon >> k) & 1) && (c[k] < S.zlo[k] || c[k] > S.zhi[k])) d = 1;
    }
    dead[i] = d;
    if (d) atomicAdd(&count[sc], 1);
}

// dead flags in rank order for the scene occupying slots b..e-1
__global__ void k_zoi_rank(Geo g, int b, int e, const unsigned char *dead, unsigned *dead_by_rank)
{
    const int i = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < e) dead_by_rank[b + g.oid[i]] = dead[i];
}

// Ranks close up over the erased spheres (their DO ids close up: DOWorld::
// EraseDOStatus); an erased sphere takes OID_DEAD | its rank in the dropped
// region.  pre: exclusive prefix count of erased spheres in rank order.
__global__ void k_zoi_newoid(Geo g, int b, int e, const unsigned char *dead, const unsigned *pre, unsigned dead_base)
{
    const int i = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e) return;
    const unsigned o = g.oid[i];
    const unsigned before = pre[b + o] - pre[b];
    g.oid[i] = dead[i] ? (OID_DEAD | (dead_base + before)) : o - before;
}

// IactRecordTab::ModifyPair (IactRecordTab.cpp:41-69) of the records: each key
// moves by DO id -> jump[id] (the new id of an erased DO is its predecessor's);
// moved keys are written over the table in key order, so of equal new keys
// the moved one wins, and of moved ones the later in old-key order.
__global__ void k_zoi_records(const ActiveOut *in, long long n, const unsigned *jump, unsigned long long *okey,
                              unsigned long long *nkey, unsigned *prio_src, RestoreRec *recs)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const ActiveOut &A = in[k];
    okey[k] = ((unsigned long long)A.lo << 32) | A.hi;
    nkey[k] = ((unsigned long long)jump[A.lo] << 32) | jump[A.hi];
    prio_src[k] = (unsigned)k;
    RestoreRec &R = recs[k];
    R.flags = A.flags & F_STATUS;
    R.pad = 0;
    R.kn = A.hist[3];
    R.init_vel = A.hist[3];
    R.shear[0] = A.hist[0]; R.shear[1] = A.hist[1]; R.shear[2] = A.hist[2];
}

__global__ void k_iota(unsigned *p, long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = (unsigned)k;
}

__global__ void k_gather_u64(const unsigned long long *src, const unsigned *idx, long long n, unsigned long long *dst)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[idx[k]];
}

__global__ void k_gather_restore(const RestoreRec *src, const unsigned *idx, long long n, RestoreRec *dst)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[idx[k]];
}

// priority of record idx_sorted[t] (t: its rank in old-key order): 0 if its
// key does not move, else 1 + t
__global__ void k_zoi_prio(const unsigned long long *okey_sorted, const unsigned *idx_sorted,
                           const unsigned long long *nkey, long long n, unsigned *prio)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const unsigned k = idx_sorted[t];
    prio[k] = nkey[k] == okey_sorted[t] ? 0u : (unsigned)(t + 1);
}

// Fresh lists for newly placed spheres (all statuses ImpactStatus()).
__global__ void k_hdr_zero(unsigned char *hdr, int b, int e)
{
    const int i = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < e) hdr[i] = 0;
}

// ---------------------------------------------------------------------------
// Export
// ---------------------------------------------------------------------------

__global__ void k_scatter_state(Geo g, int b, int e, int ld, const double *rec_imp, double *out)
{
    // out: [n][21] in the scene's sphere-rank order: pos vel ox oz av impact ang_impact
    const int i = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e) return;
    const double4 p = g.pos[i], v = g.vel[i], w = g.avl[i];
    double *d = out + (size_t)g.oid[i] * 21;
    d[0] = p.x; d[1] = p.y; d[2] = p.z;
    d[3] = v.x; d[4] = v.y; d[5] = v.z;
    d[6] = g.ori[i]; d[7] = g.ori[ld + i]; d[8] = g.ori[2 * (size_t)ld + i];
    d[9] = g.ori[3 * (size_t)ld + i]; d[10] = g.ori[4 * (size_t)ld + i]; d[11] = g.ori[5 * (size_t)ld + i];
    d[12] = w.x; d[13] = w.y; d[14] = w.z;
    for (int c = 0; c < 6; ++c) d[15 + c] = rec_imp ? rec_imp[(size_t)c * ld + i] : 0.0;
}

// Active contacts at the current positions with their ImpactStatus.  The
// candidate list is walked, not the neighbour list: a pair may come into
// contact at the end of a step whose guard has just asked for a relist (or
// after assigned statuses moved spheres), before the next step relists.  The
// neighbour list is an ordered subsequence of the candidate list, so a merge
// walk finds each listed pair's entry (flags, history, records); a candidate
// outside the list has been solved as inactive every step: Clean status.
template <bool EXT>
__global__ void k_export_active(Geo g, Ell ell, int b, int e, int ld, int ki, const unsigned *o_cnt,
                                const unsigned *o_nbr, const WallDev *walls, const unsigned *wall_id,
                                const unsigned *sphere_id, const double *rec, ActiveOut *out, unsigned long long cap,
                                unsigned long long *counter, int backup, int no_bond, const double4 *holes,
                                const PbcDev *pbc)
{
    // spheres b..e-1 (one scene); sphere_id maps the scene's ranks to DO ids
    const int i = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e) return;
    const int n = ld;
    const unsigned hdr = ell.hdr[i];
    const int h = hdr >> 7, cnt = hdr & 127;
    const double4 p = g.pos[i];
    const d3 xi = mk(p.x, p.y, p.z);
    const unsigned ec = o_cnt[i];
    int q = 0;
    unsigned listed = cnt > 0 ? (ell.ent[eidx(h, 0, ki, n, i)] & PARTNER_MASK) : 0xffffffffu;
    for (unsigned t = 0; t < ec; ++t) {
        const unsigned j = o_nbr[(size_t)t * n + i] & PARTNER_MASK;
        int k = -1;
        unsigned ent;
        if (listed == j) {
            k = q;
            ent = ell.ent[eidx(h, q, ki, n, i)];
            ++q;
            listed = q < cnt ? (ell.ent[eidx(h, q, ki, n, i)] & PARTNER_MASK) : 0xffffffffu;
        } else {
            ent = j;
            if (j >= (unsigned)n || g.oid[i] < g.oid[j]) ent |= F_SLAVE;
        }
        if (!(ent & F_SLAVE)) continue;
        Contact c;
        unsigned other;
        if (j < (unsigned)n) {
            const double4 pj = g.pos[j];
            if constexpr (EXT) cd_ss_bc(xi, p.w, mk(pj.x, pj.y, pj.z), pj.w, pbc, c);
            else cd_ss(xi, p.w, mk(pj.x, pj.y, pj.z), pj.w, c);
            other = sphere_id[g.oid[j]];
        } else {
            cd_wall_at<EXT>(xi, p.w, walls[j - n], holes, (int)(j - n), c);
            other = wall_id[j - n];
        }
        if (!c.active) continue;
        const unsigned long long slot = atomicAdd(counter, 1ull);
        if (!out || slot >= cap) continue;
        ActiveOut &A = out[slot];
        const unsigned me = sphere_id[g.oid[i]];
        A.lo = me < other ? me : other;
        A.hi = me < other ? other : me;
        // after a plain step the reference has run BackupImpactStatus (the
        // entry holds it as pending); records taken for a rebuild or a record
        // are read before it, as RebuildIactRecordTab reads them
        // (LeapConsultant.cpp:62-83)
        A.flags = (backup ? backed_up(ent & ~PARTNER_MASK) : (ent & ~PARTNER_MASK)) & ~F_NOBACKUP;
        Hist hs;
        hs.s = mk(0, 0, 0);
        hs.aux = 0.0;
        if (k >= 0 && (ent & F_CONTACT)) load_hist(ell, ell.slot[eidx(h, k, ki, n, i)], n, i, hs);
        A.hist[0] = hs.s.x; A.hist[1] = hs.s.y; A.hist[2] = hs.s.z; A.hist[3] = hs.aux;
        const bool stored = rec && k >= 0 && ((ent & F_CONTACT) || !no_bond);
        for (int r = 0; r < 13; ++r) A.rec[r] = stored ? rec[((size_t)k * 13 + r) * n + i] : 0.0;
    }
}

template __global__ void k_export_active<false>(Geo, Ell, int, int, int, int, const unsigned *, const unsigned *,
                                               const WallDev *, const unsigned *, const unsigned *, const double *,
                                               ActiveOut *, unsigned long long, unsigned long long *, int, int,
                                               const double4 *, const PbcDev *);
template __global__ void k_export_active<true>(Geo, Ell, int, int, int, int, const unsigned *, const unsigned *,
                                              const WallDev *, const unsigned *, const unsigned *, const double *,
                                              ActiveOut *, unsigned long long, unsigned long long *, int, int,
                                              const double4 *, const PbcDev *);

// Impulse on each wall from the last recorded step: sum of ImpactToMaster
// and AngularImpactToMaster over its contacts (the DOStatus impact a wall's
// ClearImpact stores), one block per wall, fixed-order reduction.
__global__ void k_wall_impulse(Ell ell, int b, int e, int ld, int ki, int w0, const double *rec, int no_bond,
                               double *out)
{
    // one block per wall w0 + blockIdx.x, over the scene's spheres b..e-1
    const int w = w0 + blockIdx.x;
    __shared__ double sm[6][256];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int i = b + threadIdx.x; i < e; i += blockDim.x) {
        const unsigned hdr = ell.hdr[i];
        const int h = hdr >> 7, cnt = hdr & 127;
        for (int k = 0; k < cnt; ++k) {
            const unsigned ent = ell.ent[eidx(h, k, ki, ld, i)];
            if ((ent & PARTNER_MASK) != (unsigned)(ld + w)) continue;
            if (!(ent & F_CONTACT) && no_bond) continue;   // a Clean status: no record stored, zero
            for (int c = 0; c < 6; ++c) acc[c] += rec[((size_t)k * 13 + 6 + c) * ld + i];
        }
    }
    for (int c = 0; c < 6; ++c) sm[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 6; ++c) sm[c][threadIdx.x] += sm[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c) out[blockIdx.x * 6 + c] = sm[c][0];
}

// Assigned sphere statuses of SimMediator::Run(map) (SimMediator.cpp:881-889):
// after Response, the listed DOs take the given status, which carries zero
// stored impulses.  buf holds 16 doubles per sphere rank: pos, vel, Ox, Oz,
// w, and a nonzero last value for the listed ones.
__global__ void k_assign_spheres(Geo g, int b, int e, int ld, const double *buf, double *rec_imp)
{
    const int i = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e) return;
    const double *d = buf + (size_t)g.oid[i] * 16;
    if (d[15] == 0.0) return;
    g.pos[i] = make_double4(d[0], d[1], d[2], g.pos[i].w);
    g.vel[i] = make_double4(d[3], d[4], d[5], 0.0);   // the host forces a relist
    for (int c = 0; c < 6; ++c) g.ori[(size_t)c * ld + i] = d[6 + c];
    g.avl[i] = make_double4(d[12], d[13], d[14], g.avl[i].w);
    if (rec_imp)
        for (int c = 0; c < 6; ++c) rec_imp[(size_t)c * ld + i] = 0.0;
}

// DOWorld::CalculateSystemEnergy (DOWorld.cpp:481-632) over the spheres (the
// mobile DOs of this path).  Every per-sphere term is formed as the reference
// forms it; the sums are a fixed-shape tree, so they differ from the
// reference's sequential sums only in rounding order.
//
// ---------------------------------------------------------------------------
// Knight analyses (src/vedo/knight) of a scene's current state
// ---------------------------------------------------------------------------

// A sphere counts for a box when its centre lies within the box widened by
// its radius, every comparison strict (CalcCoordinateNum.cpp:24-49,
// CalcDensity.cpp:20-43).
__device__ inline bool knight_in_box(const KnightBox &B, double4 p)
{
    const double r = p.w;
    return p.x < B.hi[0] + r && p.x > B.lo[0] - r && p.y < B.hi[1] + r && p.y > B.lo[1] - r && p.z < B.hi[2] + r &&
           p.z > B.lo[2] - r;
}

// CalcCoordinateNum::computeAvgCoordinateNum (CalcCoordinateNum.cpp:8-77):
// per sphere of the box, the spheres whose centre distance is at most the sum
// of the radii.  Touching spheres are candidates of each other (the candidate
// list covers every pair that can touch before the next rebuild), so each
// sphere walks its candidate list.  acc[0] += touching partners, acc[1] += 1.
__global__ void k_coordination(Geo g, int b, int e, int ld, const unsigned *o_cnt, const unsigned *o_nbr, KnightBox box,
                               unsigned long long *acc)
{
    const int i = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e) return;
    const double4 p = g.pos[i];
    if (!knight_in_box(box, p)) return;
    const d3 xi = ld3(p);
    unsigned count = 0;
    const unsigned c = o_cnt[i];
    for (unsigned t = 0; t < c; ++t) {
        const unsigned j = o_nbr[(size_t)t * ld + i] & PARTNER_MASK;
        if (j >= (unsigned)ld) continue;   // a wall
        const double4 q = g.pos[j];
        if (len(xi - ld3(q)) <= p.w + q.w) ++count;
    }
    atomicAdd(acc, (unsigned long long)count);
    atomicAdd(acc + 1, 1ull);
}

// CalcDensity::computeDensity (CalcDensity.cpp:12-72), the hit test of each
// random point: is it inside a sphere of the box?  The spheres come from the
// 27 cells around the point in the grid of the last rebuild: a sphere that
// holds the point was, at the rebuild, within its radius plus the path the
// cell size allows for (the safe distance's own assumption) of the point.
__global__ void k_density_hits(Geo g, GridDev G, long long cell_off, const int *cstart, const int *cend,
                               const double *pts, int np, KnightBox box, unsigned long long *hits)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const d3 p = mk(pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]);
    const long long cx = (long long)floor((p.x - G.xmin) / G.zr);
    const long long cy = (long long)floor((p.y - G.ymin) / G.zr);
    const long long cz = (long long)floor((p.z - G.zmin) / G.zr);
    // sphere cells run 0 .. n-3 per axis (the grid keeps a layer of padding)
    const long long x0 = max(cx - 1, 0ll), x1 = min(cx + 1, (long long)G.nx - 3);
    const long long y0 = max(cy - 1, 0ll), y1 = min(cy + 1, (long long)G.ny - 3);
    const long long z0 = max(cz - 1, 0ll), z1 = min(cz + 1, (long long)G.nz - 3);
    for (long long z = z0; z <= z1; ++z)
        for (long long y = y0; y <= y1; ++y)
            for (long long x = x0; x <= x1; ++x) {
                const long long l = cell_off + lin_of(G, x, y, z);
                const int s1 = cend[l];
                for (int s = cstart[l]; s < s1; ++s) {
                    const double4 sp = g.pos[s];
                    if (knight_in_box(box, sp) && len(p - ld3(sp)) <= sp.w) {
                        atomicAdd(hits, 1ull);
                        return;
                    }
                }
            }
}

// CalculateGranularTemperature (Knight.cpp:1162-1225) over a GSRectangle
// (GSRectangle.cpp:15-30): pass 0 writes per block the sums of v, w and the
// count of the selected spheres; pass 1 the sums of |v - avg v|^2 and
// |w - avg w|^2 (avg: 6 values).  Blocks reduce in a fixed order.
__global__ __launch_bounds__(256) void k_gt_part(Geo g, int b, int e, GtRegion q, int pass, const double *avg,
                                                 double *part)
{
    __shared__ double sm[7][256];
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = b + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
        if (g.mdl[i] != q.model) continue;
        const double4 p = g.pos[i];
        const d3 lp = ld3(p) - q.c;
        const double lx = dot(lp, q.ox), ly = dot(lp, q.oy), lz = dot(lp, q.oz);
        if (!(fabs(lx) <= 0.5 * q.w && fabs(ly) <= 0.5 * q.l && fabs(lz) <= 0.5 * q.ht)) continue;
        const double4 u = g.vel[i], w = g.avl[i];
        if (pass == 0) {
            v[0] += u.x; v[1] += u.y; v[2] += u.z;
            v[3] += w.x; v[4] += w.y; v[5] += w.z;
            v[6] += 1.0;
        } else {
            const d3 dv = ld3(u) - mk(avg[0], avg[1], avg[2]), dw = ld3(w) - mk(avg[3], avg[4], avg[5]);
            v[0] += dot(dv, dv);
            v[1] += dot(dw, dw);
        }
    }
    for (int k = 0; k < 7; ++k) sm[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
        if ((int)threadIdx.x < s2)
            for (int k = 0; k < 7; ++k) sm[k][threadIdx.x] += sm[k][threadIdx.x + s2];
        __syncthreads();
    }
    if (threadIdx.x < 7) part[(size_t)blockIdx.x * 7 + threadIdx.x] = sm[threadIdx.x][0];
}

// k_energy_part, pass 0, writes per block: [0] sum m (g.x), [1] sum 0.5 m v v,
// [2] sum 0.5 (I wx wx + I wy wy + I wz wz), [3..5] sum m v, [6..8] sum
// mmi x w (the reference's "angular momentum"), [9] max |v|, [10] min |v|,
// [11] max |w|, [12] min |w|.  Pass 1 writes [0] sum |m v - P| and
// [1] sum |mmi x w - L| for the averages P, L in avg[0..5].
static constexpr int EN_K0 = 13, EN_K1 = 2, EN_BLOCK = 256;

__device__ inline void en_terms(const Geo &g, int i, const double2 *mm, double gx, double gy, double gz, d3 &mv,
                                d3 &Lw, double *t)
{
    const double4 p = g.pos[i], v4 = g.vel[i], w4 = g.avl[i];
    const double2 M = mm[g.mdl[i]];
    const double m = M.x, I = M.y;
    const d3 vel = ld3(v4), av = ld3(w4);
    const double v = len(vel);
    mv = smul(m, vel);
    Lw = cross(mk(I, I, I), av);
    t[0] = m * (((gx * p.x) + (gy * p.y)) + (gz * p.z));
    t[1] = 0.5 * m * v * v;
    t[2] = 0.5 * ((((I * av.x) * av.x) + ((I * av.y) * av.y)) + ((I * av.z) * av.z));
    t[3] = v;
    t[4] = len(av);
}

template <int K>
__device__ inline void en_block_reduce(double (*sh)[EN_BLOCK], double *acc, int nsum, double *out)
{
    const int t = threadIdx.x;
    for (int k = 0; k < K; ++k) sh[k][t] = acc[k];
    __syncthreads();
    for (int s = EN_BLOCK / 2; s > 0; s >>= 1) {
        if (t < s)
            for (int k = 0; k < K; ++k) {
                const double a = sh[k][t], b = sh[k][t + s];
                sh[k][t] = k < nsum ? a + b : (((k - nsum) & 1) ? fmin(a, b) : fmax(a, b));
            }
        __syncthreads();
    }
    if (t < K) out[t] = sh[t][0];
}

__global__ __launch_bounds__(EN_BLOCK) void k_energy_part(Geo g, int b, int e, const double2 *mm, double gx,
                                                          double gy, double gz, const double *avg, int pass,
                                                          double *part)
{
    __shared__ double sh[EN_K0][EN_BLOCK];
    double acc[EN_K0];
    for (int k = 0; k < EN_K0; ++k) acc[k] = 0.0;
    acc[9] = acc[11] = -INFINITY;
    acc[10] = acc[12] = INFINITY;
    for (int i = b + blockIdx.x * EN_BLOCK + threadIdx.x; i < e; i += gridDim.x * EN_BLOCK) {
        d3 mv, Lw;
        double t[5];
        en_terms(g, i, mm, gx, gy, gz, mv, Lw, t);
        if (pass == 0) {
            acc[0] += t[0]; acc[1] += t[1]; acc[2] += t[2];
            acc[3] += mv.x; acc[4] += mv.y; acc[5] += mv.z;
            acc[6] += Lw.x; acc[7] += Lw.y; acc[8] += Lw.z;
            acc[9] = fmax(acc[9], t[3]); acc[10] = fmin(acc[10], t[3]);
            acc[11] = fmax(acc[11], t[4]); acc[12] = fmin(acc[12], t[4]);
        } else {
            acc[0] += len(mv - mk(avg[0], avg[1], avg[2]));
            acc[1] += len(Lw - mk(avg[3], avg[4], avg[5]));
        }
    }
    if (pass == 0) en_block_reduce<EN_K0>(sh, acc, 9, part + (size_t)blockIdx.x * EN_K0);
    else en_block_reduce<EN_K1>(sh, acc, 2, part + (size_t)blockIdx.x * EN_K1);
}

// One block: folds the per-block partials in block order and writes the
// results (out layout of vg_system_energy).  Pass 0 also leaves the averages
// for pass 1 in avg[0..5].
__global__ __launch_bounds__(EN_BLOCK) void k_energy_final(const double *part, int nb, int n, int pass, double *avg,
                                                           double *out)
{
    __shared__ double sh[EN_K0][EN_BLOCK];
    __shared__ double r[EN_K0];
    const int K = pass == 0 ? EN_K0 : EN_K1, nsum = pass == 0 ? 9 : 2;
    double acc[EN_K0];
    for (int k = 0; k < EN_K0; ++k) acc[k] = 0.0;
    acc[9] = acc[11] = -INFINITY;
    acc[10] = acc[12] = INFINITY;
    for (int b = threadIdx.x; b < nb; b += EN_BLOCK)
        for (int k = 0; k < K; ++k) {
            const double x = part[(size_t)b * K + k];
            acc[k] = k < nsum ? acc[k] + x : (((k - nsum) & 1) ? fmin(acc[k], x) : fmax(acc[k], x));
        }
    if (pass == 0) en_block_reduce<EN_K0>(sh, acc, 9, r);
    else en_block_reduce<EN_K1>(sh, acc, 2, r);
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (pass == 0) {
        // vMomentumAvg = 1.0 / NumMobileDO * vMomentumAvg (DOWorld.cpp:589-590)
        const double f = 1.0 / (double)n;
        for (int c = 0; c < 6; ++c) avg[c] = f * r[3 + c];
        out[0] = n ? -r[0] : 0.0;
        out[1] = r[1]; out[2] = r[2];
        for (int c = 0; c < 6; ++c) out[3 + c] = avg[c];
        out[11] = n ? r[9] : 0.0; out[12] = n ? r[10] : 0.0;
        out[13] = n ? r[11] : 0.0; out[14] = n ? r[12] : 0.0;
        out[15] = (double)n;
    } else {
        out[9] = r[0]; out[10] = r[1];
    }
}

// Plain streaming copy used to calibrate HBM counters and measure achievable
// bandwidth.
// Streaming copy (the achievable HBM bandwidth): each workgroup copies
// chunks of 4 x blockDim 16-byte elements, the four loads of a thread in
// flight together, streaming loads and stores.
__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n)
{
    constexpr int U = 4;
    const size_t chunk = (size_t)blockDim.x * U;
    const v2d *s = reinterpret_cast<const v2d *>(src);
    v2d *d = reinterpret_cast<v2d *>(dst);
    for (size_t base = (size_t)blockIdx.x * chunk; base < n; base += (size_t)gridDim.x * chunk) {
        v2d v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = base + (size_t)u * blockDim.x + threadIdx.x;
            if (i < n) v[u] = __builtin_nontemporal_load(s + i);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = base + (size_t)u * blockDim.x + threadIdx.x;
            if (i < n) __builtin_nontemporal_store(v[u], d + i);
        }
    }
}

}  // namespace vg
