// vg_api.hip -- host side of libvedo_hip.so: the C-ABI of include/vedo_gpu.h.
//
// Plays the role of VEDO's SimMediator (SimMediator.cpp:640-842) and
// LeapConsultant (LeapConsultant.cpp:49-195) for one scene on one GPU:
// holds the device state, orders the per-step kernels on its own stream, and
// runs the candidate-list rebuild every up_iact steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include <rocprim/rocprim.hpp>
#include <string>
#include <vector>

#include "../../include/vedo_gpu.h"
#include "vg_kernels.h"
// one translation unit: the kernels are launched from this file
#include "vg_kernels.hip"


using namespace vg;

static std::string g_create_error;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct vg_scene {
    std::string err;
    hipStream_t stream = nullptr;
    int device = 0;
    // sizes
    int n = 0, nw = 0, n_do = 0, ki = 48, n_groups = 0;
    // host mirrors
    std::vector<int> do_is_sphere;         // DO id -> 1 sphere / 0 wall
    std::vector<unsigned> wall_ids;        // wall index -> DO id
    std::vector<unsigned> sphere_ids;      // sphere rank -> DO id
    std::vector<int> do_to_sphere;         // DO id -> sphere rank or -1
    unsigned *d_sphere_ids = nullptr, *d_wall_ids = nullptr;
    std::vector<int> do_to_wall;           // DO id -> wall index or -1
    std::vector<WallDev> walls_h;
    std::vector<ModelDev> models_h;
    std::vector<LawDev> laws_h;
    bool sd_only = false;      // every law is of the spring-dashpot family
    int single_group = 0, no_bond = 1;
    // scalars
    double dt = 0, t = 0, t_stop = 0, g[3] = {0, 0, 0}, sf = 1.1, skin = 0, bond_reach = 0;
    unsigned up_iact = 1, record = 0;
    unsigned long long round = 0, step_count = 0;
    long long rebuilds = 0;
    long long n_iact = 0;
    // device
    Geo geo[2];
    int cur = 0;
    Ell ell[2];
    int ecur = 0;
    ModelDev *models = nullptr;
    LawDev *laws = nullptr;
    WallDev *walls = nullptr, *walls_ref = nullptr;
    WallKin *wkin = nullptr;
    double *assigned = nullptr;
    double *assigned_sph = nullptr;   // [n][16] assigned sphere statuses by rank (allocated on first use)
    double2 *mm = nullptr;            // per model: mass, mass moment of inertia (spheres)
    double *en_part = nullptr, *en_out = nullptr;
    int *flags = nullptr;
    unsigned long long *relists = nullptr;
    unsigned long long *red = nullptr, *bpart = nullptr;
    unsigned *inv = nullptr, *perm = nullptr, *svals = nullptr;
    unsigned long long *keys = nullptr, *skeys = nullptr;
    double *xs[2] = {nullptr, nullptr}, *ys[2] = {nullptr, nullptr}, *zs[2] = {nullptr, nullptr};
    double *safe = nullptr;
    int *lin = nullptr;
    int *cstart = nullptr, *cend = nullptr;
    long long cells_cap = 0;
    unsigned *ocount = nullptr, *onbr = nullptr;
    int ko = 64;   // candidate-list capacity per sphere (grows)
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    double *rec = nullptr, *rec_imp = nullptr, *rec_imp2 = nullptr;
    bool rec_valid = false;   // rec_imp holds the last step's impulses
    bool ent_valid = false;   // rec (per-entry ImpactStatus fields) matches the current lists
    // IactRecordTab snapshot taken before a rebuild re-sorts the spheres
    ActiveOut *snap = nullptr;
    size_t snap_cap = 0;
    long long snap_count = 0;
    double *snap_wimp = nullptr;
    bool snap_valid = false;
    // timing
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;
    int last_launches = 0;
    bool timing = false;
    // per-launch events around the step kernel (vg_kernel_timing)
    std::vector<hipEvent_t> kev;
    int kev_used = 0;
};

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            h->err = std::string(#x) + ": " + hipGetErrorString(e_);                           \
            return VG_EHIP;                                                                    \
        }                                                                                      \
    } while (0)

template <class T> static int dalloc(vg_scene *h, T **p, size_t count)
{
    if (count == 0) count = 1;
    HIPCHK(hipMalloc((void **)p, sizeof(T) * count));
    return 0;
}

static int fail(vg_scene *h, int code, const std::string &msg)
{
    h->err = msg;
    return code;
}

static inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

static double olen(const double *v) { return std::sqrt((v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2])); }

static double from_ord(unsigned long long u)
{
    unsigned long long b = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}

static int ensure_tmp(vg_scene *h, size_t bytes)
{
    if (bytes <= h->tmp_bytes) return 0;
    if (h->tmp) HIPCHK(hipFree(h->tmp));
    h->tmp_bytes = bytes + bytes / 4 + 1024;
    HIPCHK(hipMalloc(&h->tmp, h->tmp_bytes));
    return 0;
}

static Tables tables(vg_scene *h) { return Tables{h->models, h->laws, h->n_groups}; }

// LeapConsultant::Reset + subReset (LeapConsultant.cpp:108-195,
// LeapConsultant.h:57-514) followed by SimMediator::Initiate's history restore.
static int rebuild(vg_scene *h)
{
    const int n = h->n;
    hipStream_t st = h->stream;
    Geo &G0 = h->geo[h->cur];
    Geo &G1 = h->geo[h->cur ^ 1];
    // bounds
    const int nb = (int)nblk(n, 256);
    hipLaunchKernelGGL(k_bounds, dim3(nb), dim3(256), 0, st, G0, n, h->bpart, h->inv, h->xs[0], h->ys[0],
                       h->zs[0]);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, st, (const unsigned long long *)h->bpart, nb, h->red);
    // 2%-trimmed spans pick the comparator (LeapConsultant.cpp:139-191)
    size_t need = 0;
    (void)rocprim::radix_sort_keys(nullptr, need, h->xs[0], h->xs[1], (size_t)n, 0, 64, st);
    if (ensure_tmp(h, need)) return VG_EHIP;
    HIPCHK(rocprim::radix_sort_keys(h->tmp, need, h->xs[0], h->xs[1], (size_t)n, 0, 64, st));
    HIPCHK(rocprim::radix_sort_keys(h->tmp, need, h->ys[0], h->ys[1], (size_t)n, 0, 64, st));
    HIPCHK(rocprim::radix_sort_keys(h->tmp, need, h->zs[0], h->zs[1], (size_t)n, 0, 64, st));
    unsigned long long red[8];
    HIPCHK(hipMemcpyAsync(red, h->red, sizeof(red), hipMemcpyDeviceToHost, st));
    const double discard = 0.02, keep = 1.0 - discard;
    const size_t ihi = (size_t)(keep * n - 1), ilo = (size_t)(discard * n);
    double q[6];
    HIPCHK(hipMemcpyAsync(&q[0], h->xs[1] + ilo, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&q[1], h->xs[1] + ihi, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&q[2], h->ys[1] + ilo, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&q[3], h->ys[1] + ihi, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&q[4], h->zs[1] + ilo, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&q[5], h->zs[1] + ihi, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const double xmax = from_ord(red[0]), ymax = from_ord(red[1]), zmax = from_ord(red[2]);
    const double xmin = from_ord(~red[3]), ymin = from_ord(~red[4]), zmin = from_ord(~red[5]);
    const double vmax = from_ord(red[6]), rmax = from_ord(red[7]);
    const double xspan = q[1] - q[0], yspan = q[3] - q[2], zspan = q[5] - q[4];
    GridDev gd;
    // TrirLeX (x,y,z), TrirLeY (y,x,z), TrirLeZ (z,x,y) (NBSConsultant.h:52-118)
    int perm[3];
    if (xspan >= yspan) {
        if (xspan >= zspan) { perm[0] = 0; perm[1] = 1; perm[2] = 2; }
        else { perm[0] = 2; perm[1] = 0; perm[2] = 1; }
    } else {
        if (yspan >= zspan) { perm[0] = 1; perm[1] = 0; perm[2] = 2; }
        else { perm[0] = 2; perm[1] = 0; perm[2] = 1; }
    }
    const double dtu = h->up_iact * h->dt;
    const double glen = olen(h->g);
    const double zr = h->sf * (2.0 * rmax + dtu * vmax + 0.5 * dtu * dtu * glen);
    unsigned ncx = (unsigned)std::ceil((xmax - xmin) / zr); ncx = std::max(ncx, 1u);
    unsigned ncy = (unsigned)std::ceil((ymax - ymin) / zr); ncy = std::max(ncy, 1u);
    unsigned ncz = (unsigned)std::ceil((zmax - zmin) / zr); ncz = std::max(ncz, 1u);
    if (ncx >= (1u << 21) || ncy >= (1u << 21) || ncz >= (1u << 21))
        return fail(h, VG_EUNSUPPORTED, "cell grid exceeds 2^21 cells per axis");
    gd.xmin = xmin; gd.ymin = ymin; gd.zmin = zmin; gd.zr = zr;
    gd.safe_v = dtu;
    gd.safe_g = 0.5 * dtu * dtu * glen;
    gd.sf = h->sf;
    gd.nx = (int)ncx + 3; gd.ny = (int)ncy + 3; gd.nz = (int)ncz + 3;
    gd.perm[0] = perm[0]; gd.perm[1] = perm[1]; gd.perm[2] = perm[2];
    const long long cells = (long long)gd.nx * gd.ny * gd.nz;
    if (cells > (1ll << 31)) return fail(h, VG_EUNSUPPORTED, "cell grid too large (sparse domain)");
    if (cells > h->cells_cap) {
        if (h->cstart) HIPCHK(hipFree(h->cstart));
        if (h->cend) HIPCHK(hipFree(h->cend));
        h->cells_cap = cells + cells / 4;
        if (dalloc(h, &h->cstart, h->cells_cap) || dalloc(h, &h->cend, h->cells_cap)) return VG_ENOMEM;
    }
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(cells, 256)), dim3(256), 0, st, h->cstart, cells, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(cells, 256)), dim3(256), 0, st, h->cend, cells, 0);
    // stable sort by (comparator cell key), DO-id order inside a cell
    hipLaunchKernelGGL(k_cellkey, dim3(nblk(n, 256)), dim3(256), 0, st, G0, n, gd, h->inv, h->keys, h->svals);
    need = 0;
    (void)rocprim::radix_sort_pairs(nullptr, need, h->keys, h->skeys, h->svals, h->perm, (size_t)n, 0, 63, st);
    if (ensure_tmp(h, need)) return VG_EHIP;
    HIPCHK(rocprim::radix_sort_pairs(h->tmp, need, h->keys, h->skeys, h->svals, h->perm, (size_t)n, 0, 63, st));
    hipLaunchKernelGGL(k_reorder, dim3(nblk(n, 256)), dim3(256), 0, st, G0, G1, n, h->perm, gd, h->safe, h->lin,
                       h->rec_valid ? (const double *)h->rec_imp : (const double *)nullptr, h->rec_imp2);
    if (h->rec_valid) std::swap(h->rec_imp, h->rec_imp2);
    hipLaunchKernelGGL(k_cellstart, dim3(nblk(n, 256)), dim3(256), 0, st, h->lin, n, h->cstart, h->cend);
    // candidate list in one pass; grow the capacity and redo if a sphere overflowed
    OuterArgs oa;
    oa.g = G1; oa.n = n; oa.nw = h->nw; oa.G = gd; oa.cstart = h->cstart; oa.cend = h->cend; oa.safe = h->safe;
    oa.walls = h->walls; oa.t = tables(h); oa.count = h->ocount; oa.nbr = h->onbr; oa.ko = h->ko;
    oa.single_group = h->single_group; oa.flags = h->flags;
    for (;;) {
        HIPCHK(hipMemsetAsync(h->flags + 5, 0, sizeof(int), st));
        hipLaunchKernelGGL(k_outer, dim3(nblk(n, 128)), dim3(128), 0, st, oa);
        int maxc = 0;
        HIPCHK(hipMemcpyAsync(&maxc, h->flags + 5, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (maxc <= h->ko) break;
        HIPCHK(hipFree(h->onbr));
        h->onbr = nullptr;
        h->ko = maxc + maxc / 4 + 8;
        if (dalloc(h, &h->onbr, (size_t)h->ko * n)) return VG_ENOMEM;
        oa.nbr = h->onbr;
        oa.ko = h->ko;
    }
    h->n_iact = -1;   // counted on request (vg_num_interactions)
    // neighbour lists with history carry
    InnerArgs ia;
    ia.g = G1; ia.gold = G0; ia.ell = h->ell[h->ecur ^ 1]; ia.eold = h->ell[h->ecur];
    ia.n = n; ia.nw = h->nw; ia.ki = h->ki; ia.perm = h->perm; ia.cnt = h->ocount; ia.nbr = h->onbr; ia.ko = h->ko;
    ia.walls = h->walls; ia.skin = h->skin; ia.flags = h->flags;
    hipLaunchKernelGGL(k_inner_build, dim3(nblk(n, 128)), dim3(128), (size_t)h->ki * 128 * sizeof(unsigned), st,
                       ia);
    // the fresh lists need no relist: clear the ring slot the next step reads
    // and restart the wall displacement bounds
    HIPCHK(hipMemsetAsync(h->flags + ((h->step_count) % 3), 0, sizeof(int), st));
    if (h->nw) hipLaunchKernelGGL(k_wall_path_reset, dim3(nblk(h->nw, 64)), dim3(64), 0, st, h->walls, h->nw);
    HIPCHK(hipGetLastError());
    h->cur ^= 1;
    h->ecur ^= 1;
    h->ent_valid = false;
    h->rebuilds++;
    return 0;
}

static void set_err_create(const std::string &s) { g_create_error = s; }

extern "C" const char *vg_create_error(void) { return g_create_error.c_str(); }

extern "C" const char *vg_last_error(const vg_scene *h) { return h ? h->err.c_str() : g_create_error.c_str(); }

static int alloc_geo(vg_scene *h, Geo &g, int n)
{
    if (dalloc(h, &g.pos, n) || dalloc(h, &g.vel, n) || dalloc(h, &g.avl, n) || dalloc(h, &g.ori, 6 * (size_t)n) ||
        dalloc(h, &g.oid, n) || dalloc(h, &g.mdl, n))
        return VG_ENOMEM;
    return 0;
}

static int alloc_ell(vg_scene *h, Ell &e, int n, int ki)
{
    if (dalloc(h, &e.hdr, n) || dalloc(h, &e.ent, (size_t)2 * ki * n) || dalloc(h, &e.hist, (size_t)2 * ki * n))
        return VG_ENOMEM;
    return 0;
}

static void free_scene(vg_scene *h)
{
    auto F = [](void *p) {
        if (p) (void)hipFree(p);
    };
    for (int b = 0; b < 2; ++b) {
        F(h->geo[b].pos); F(h->geo[b].vel); F(h->geo[b].avl); F(h->geo[b].ori); F(h->geo[b].oid); F(h->geo[b].mdl);
        F(h->ell[b].hdr); F(h->ell[b].ent); F(h->ell[b].hist);
        F(h->xs[b]); F(h->ys[b]); F(h->zs[b]);
    }
    F(h->models); F(h->laws); F(h->walls); F(h->walls_ref); F(h->wkin); F(h->assigned); F(h->flags);
    F(h->relists); F(h->red); F(h->bpart); F(h->inv); F(h->perm); F(h->svals); F(h->keys); F(h->skeys); F(h->safe); F(h->lin);
    F(h->cstart); F(h->cend); F(h->ocount); F(h->onbr); F(h->tmp); F(h->rec); F(h->rec_imp); F(h->rec_imp2); F(h->snap); F(h->snap_wimp); F(h->d_sphere_ids); F(h->d_wall_ids);
    F(h->assigned_sph); F(h->mm); F(h->en_part); F(h->en_out);
    for (auto e : h->kev) (void)hipEventDestroy(e);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

extern "C" int vg_destroy(vg_scene *h)
{
    if (!h) return VG_EINVAL;
    free_scene(h);
    delete h;
    return 0;
}

static double model_range(const vg_model &m)
{
    // DOModel::Analysis (DOModel.cpp:17-149), global walls
    if (m.shape == VG_SPHERE) return m.attr[0];
    if (m.shape == VG_QUASICYLINDER) return m.scope == VG_LOCAL ? std::max(m.attr[0], 0.5 * m.attr[1]) : m.attr[0];
    if (m.shape == VG_QUASIPLATE) {
        if (m.scope == VG_LOCAL) return 0.5 * std::max(std::max(m.attr[0], m.attr[1]), m.attr[2]);
        return 0.5 * m.attr[1];
    }
    return 0.0;
}

extern "C" int vg_create(const vg_scene_desc *d, vg_scene **out)
{
    if (!d || !out) {
        set_err_create("null argument");
        return VG_EINVAL;
    }
    *out = nullptr;
    vg_scene *h = new vg_scene();
    auto bail = [&](int code) {
        set_err_create(h->err);
        free_scene(h);
        delete h;
        return code;
    };
    if (d->consultant != 0) {
        h->err = "only LeapConsultant candidate-list rules run on the GPU";
        return bail(VG_EUNSUPPORTED);
    }
    const double PI = 4.0 * std::atan(1.0);
    h->n_do = d->n_do;
    h->n_groups = d->n_groups;
    h->do_is_sphere.assign(d->n_do, 0);
    h->do_to_wall.assign(d->n_do, -1);
    std::vector<unsigned> &sphere_ids = h->sphere_ids;
    for (int i = 0; i < d->n_do; ++i) {
        const int m = d->model_of[i];
        if (m < 0 || m >= d->n_models) {
            h->err = "DOStatus names an unknown DOModel";
            return bail(VG_EINVAL);
        }
        const vg_model &M = d->models[m];
        if (M.density < 0.0) {
            h->err = "negative density";
            return bail(VG_EUNSUPPORTED);
        }
        if (M.shape == VG_SPHERE) {
            if (M.scope != VG_LOCAL || M.behavior != VG_MOBILE || M.density == 0.0) {
                h->err = "spheres must be local, mobile and of positive density";
                return bail(VG_EUNSUPPORTED);
            }
            h->do_is_sphere[i] = 1;
            sphere_ids.push_back((unsigned)i);
        } else if (M.shape == VG_QUASIPLATE || M.shape == VG_QUASICYLINDER) {
            if (M.scope != VG_GLOBAL || (M.behavior != VG_FIXED && M.behavior != VG_ORBITAL)) {
                h->err = "walls must be global and fixed or orbital";
                return bail(VG_EUNSUPPORTED);
            }
            h->do_to_wall[i] = (int)h->wall_ids.size();
            h->wall_ids.push_back((unsigned)i);
        } else {
            h->err = "shape is not on the GPU path";
            return bail(VG_EUNSUPPORTED);
        }
    }
    h->n = (int)sphere_ids.size();
    h->nw = (int)h->wall_ids.size();
    if (h->n == 0) {
        h->err = "scene has no spheres";
        return bail(VG_EINVAL);
    }
    if (h->n + h->nw > (int)PARTNER_MASK) {
        h->err = "too many DOs for the neighbour-list encoding";
        return bail(VG_EUNSUPPORTED);
    }
    // tables
    h->models_h.resize(d->n_models);
    for (int m = 0; m < d->n_models; ++m) {
        h->models_h[m].group = d->models[m].group;
        for (int c = 0; c < 3; ++c) h->models_h[m].fext[c] = d->models[m].fext[c];
    }
    h->laws_h.resize((size_t)d->n_groups * d->n_groups);
    double bond_reach = 0.0;
    for (int k = 0; k < d->n_groups * d->n_groups; ++k) {
        h->laws_h[k].law = d->iact[k].law;
        for (int p = 0; p < 8; ++p) h->laws_h[k].p[p] = d->iact[k].p[p];
        if (d->iact[k].law < 0 || d->iact[k].law > 5) {
            h->err = "unknown impact solver";
            return bail(VG_EINVAL);
        }
        const int L = d->iact[k].law;
        if ((L == VG_BSDBF || L == VG_BSDB2F) && d->iact[k].p[4] != 0.0 && d->iact[k].p[0] != 0.0)
            bond_reach = std::max(bond_reach, std::fabs(d->iact[k].p[4] / d->iact[k].p[0]));
    }
    h->sd_only = true;
    for (const LawDev &L : h->laws_h) {
        if (L.law == VG_H || L.law == VG_HERTZ) h->sd_only = false;
        if ((L.law == VG_BSDBF || L.law == VG_BSDB2F || L.law == VG_BTSDBF) && L.p[4] != 0.0) h->no_bond = 0;
    }
    {
        int g0 = -1;
        h->single_group = 1;
        for (unsigned id : sphere_ids) {
            const int g = d->models[d->model_of[id]].group;
            if (g0 < 0) g0 = g;
            else if (g != g0) h->single_group = 0;
        }
    }
    h->dt = d->dt; h->t = d->t_current; h->t_stop = d->t_stop;
    for (int c = 0; c < 3; ++c) h->g[c] = d->g[c];
    h->sf = d->safety_factor; h->up_iact = d->up_iact ? d->up_iact : 1; h->record = d->record;
    double rmin = 1e300;
    for (unsigned id : sphere_ids) rmin = std::min(rmin, d->models[d->model_of[id]].attr[0]);
    h->skin = d->inner_skin > 0 ? d->inner_skin : 0.1 * rmin;
    h->skin = std::max(h->skin, bond_reach * 1.01);
    h->bond_reach = bond_reach;
    const char *ki_env = getenv("VG_NEIGHBOUR_CAPACITY");
    if (ki_env) h->ki = std::max(8, std::min(64, atoi(ki_env)));   // the step kernel keeps a 64-bit entry mask
    // walls
    h->walls_h.resize(h->nw);
    for (int w = 0; w < h->nw; ++w) {
        const unsigned id = h->wall_ids[w];
        const vg_model &M = d->models[d->model_of[id]];
        WallDev &W = h->walls_h[w];
        W.pos = mk(d->pos[3 * id], d->pos[3 * id + 1], d->pos[3 * id + 2]);
        W.vel = mk(d->vel[3 * id], d->vel[3 * id + 1], d->vel[3 * id + 2]);
        W.ox = mk(d->ox[3 * id], d->ox[3 * id + 1], d->ox[3 * id + 2]);
        W.oz = mk(d->oz[3 * id], d->oz[3 * id + 1], d->oz[3 * id + 2]);
        W.av = mk(d->av[3 * id], d->av[3 * id + 1], d->av[3 * id + 2]);
        W.a0 = M.attr[0]; W.a1 = M.attr[1]; W.a2 = M.attr[2];
        W.range = model_range(M);
        W.reach = M.shape == VG_QUASIPLATE ? 0.5 * std::sqrt(M.attr[0] * M.attr[0] + M.attr[1] * M.attr[1] +
                                                             M.attr[2] * M.attr[2])
                                           : std::sqrt(M.attr[0] * M.attr[0] + 0.25 * M.attr[1] * M.attr[1]);
        W.path = 0.0;
        W.shape = M.shape; W.group = M.group; W.behavior = M.behavior; W.kin = 0;
    }
    // device
    if (hipGetDevice(&h->device) != hipSuccess) {
        h->err = "no HIP device";
        return bail(VG_EHIP);
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        h->err = "hipStreamCreate failed";
        return bail(VG_EHIP);
    }
    (void)hipEventCreate(&h->ev0);
    (void)hipEventCreate(&h->ev1);
    const int n = h->n;
    int rc = 0;
    rc |= alloc_geo(h, h->geo[0], n) | alloc_geo(h, h->geo[1], n);
    rc |= alloc_ell(h, h->ell[0], n, h->ki) | alloc_ell(h, h->ell[1], n, h->ki);
    rc |= dalloc(h, &h->models, d->n_models) | dalloc(h, &h->laws, (size_t)d->n_groups * d->n_groups);
    rc |= dalloc(h, &h->walls, h->nw) | dalloc(h, &h->walls_ref, h->nw) | dalloc(h, &h->wkin, h->nw);
    rc |= dalloc(h, &h->assigned, (size_t)15 * h->nw);
    rc |= dalloc(h, &h->flags, 8) | dalloc(h, &h->relists, 1) | dalloc(h, &h->red, 8);
    rc |= dalloc(h, &h->bpart, (size_t)8 * nblk(n, 256));
    rc |= dalloc(h, &h->inv, n) | dalloc(h, &h->perm, n) | dalloc(h, &h->svals, n);
    rc |= dalloc(h, &h->keys, n) | dalloc(h, &h->skeys, n);
    for (int b = 0; b < 2; ++b) rc |= dalloc(h, &h->xs[b], n) | dalloc(h, &h->ys[b], n) | dalloc(h, &h->zs[b], n);
    rc |= dalloc(h, &h->safe, n) | dalloc(h, &h->lin, n);
    rc |= dalloc(h, &h->ocount, (size_t)n + 1);
    rc |= dalloc(h, &h->rec, (size_t)h->ki * 13 * n) | dalloc(h, &h->rec_imp, (size_t)6 * n);
    rc |= dalloc(h, &h->rec_imp2, (size_t)6 * n) | dalloc(h, &h->snap_wimp, (size_t)6 * std::max(h->nw, 1));
    rc |= dalloc(h, &h->d_sphere_ids, n) | dalloc(h, &h->d_wall_ids, std::max(h->nw, 1));
    rc |= dalloc(h, &h->mm, d->n_models) | dalloc(h, &h->en_out, 24);
    rc |= dalloc(h, &h->en_part, (size_t)16 * std::max(1u, std::min(1024u, nblk(n, 256))));
    if (rc) return bail(VG_ENOMEM);
    if (dalloc(h, &h->onbr, (size_t)h->ko * n)) return bail(VG_ENOMEM);
    // upload spheres in DO-id order
    {
        std::vector<double4> pos(n), vel(n), avl(n);
        std::vector<double> ori(6 * (size_t)n);
        std::vector<unsigned> oid(n);
        std::vector<int> mdl(n);
        for (int s = 0; s < n; ++s) {
            const unsigned id = sphere_ids[s];
            const vg_model &M = d->models[d->model_of[id]];
            const double r = M.attr[0];
            const double mass = M.density * ((4.0 / 3.0 * PI) * r * r * r);   // DOModel.cpp:22-27
            const double sudo = M.density_factor * mass;
            pos[s] = make_double4(d->pos[3 * id], d->pos[3 * id + 1], d->pos[3 * id + 2], r);
            vel[s] = make_double4(d->vel[3 * id], d->vel[3 * id + 1], d->vel[3 * id + 2], 0.0);
            avl[s] = make_double4(d->av[3 * id], d->av[3 * id + 1], d->av[3 * id + 2], sudo);
            for (int c = 0; c < 3; ++c) {
                ori[(size_t)c * n + s] = d->ox[3 * id + c];
                ori[(size_t)(3 + c) * n + s] = d->oz[3 * id + c];
            }
            oid[s] = (unsigned)s;
            mdl[s] = d->model_of[id];
        }
        Geo &g = h->geo[0];
        if (hipMemcpy(g.pos, pos.data(), sizeof(double4) * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(g.vel, vel.data(), sizeof(double4) * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(g.avl, avl.data(), sizeof(double4) * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(g.ori, ori.data(), sizeof(double) * 6 * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(g.oid, oid.data(), sizeof(unsigned) * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(g.mdl, mdl.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess) {
            h->err = "upload failed";
            return bail(VG_EHIP);
        }
    }
    // DOModel::Analysis mass and mass moment of inertia of spheres (DOModel.cpp:22-29)
    std::vector<double2> mm(d->n_models, make_double2(0.0, 0.0));
    for (int m = 0; m < d->n_models; ++m)
        if (d->models[m].shape == VG_SPHERE) {
            const double r = d->models[m].attr[0];
            const double mass = d->models[m].density * ((4.0 / 3.0 * PI) * r * r * r);
            mm[m] = make_double2(mass, 0.4 * mass * r * r);
        }
    h->do_to_sphere.assign(h->n_do, -1);
    for (int s = 0; s < n; ++s) h->do_to_sphere[sphere_ids[s]] = s;
    if (hipMemcpy(h->mm, mm.data(), sizeof(double2) * d->n_models, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->models, h->models_h.data(), sizeof(ModelDev) * d->n_models, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(h->laws, h->laws_h.data(), sizeof(LawDev) * h->laws_h.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        (h->nw && hipMemcpy(h->walls, h->walls_h.data(), sizeof(WallDev) * h->nw, hipMemcpyHostToDevice) !=
                      hipSuccess) ||
        hipMemcpy(h->d_sphere_ids, sphere_ids.data(), sizeof(unsigned) * n, hipMemcpyHostToDevice) != hipSuccess ||
        (h->nw && hipMemcpy(h->d_wall_ids, h->wall_ids.data(), sizeof(unsigned) * h->nw, hipMemcpyHostToDevice) !=
                      hipSuccess) ||
        hipMemset(h->flags, 0, 8 * sizeof(int)) != hipSuccess ||
        hipMemset(h->relists, 0, sizeof(unsigned long long)) != hipSuccess) {
        h->err = "upload failed";
        return bail(VG_EHIP);
    }
    hipLaunchKernelGGL(k_hdr_zero, dim3(nblk(n, 256)), dim3(256), 0, h->stream, h->ell[0].hdr, n);
    hipLaunchKernelGGL(k_hdr_zero, dim3(nblk(n, 256)), dim3(256), 0, h->stream, h->ell[1].hdr, n);
    if (rebuild(h)) return bail(VG_EHIP);
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
        h->err = "device error during the first candidate-list build";
        return bail(VG_EHIP);
    }
    *out = h;
    return 0;
}

static StepArgs step_args(vg_scene *h)
{
    StepArgs a;
    a.n = h->n; a.nw = h->nw; a.ki = h->ki;
    a.in = h->geo[h->cur]; a.out = h->geo[h->cur ^ 1];
    a.ell = h->ell[h->ecur];
    a.o_cnt = h->ocount; a.o_nbr = h->onbr; a.ko = h->ko;
    a.t = tables(h);
    a.walls = h->walls;
    a.flags = h->flags; a.relists = h->relists;
    a.dt = h->dt; a.gx = h->g[0]; a.gy = h->g[1]; a.gz = h->g[2];
    a.skin = h->skin;
    a.relist_half = 0.5 * h->skin * (1.0 - 1e-9);
    a.step = (int)(h->step_count % 3);
    a.rec_imp = h->rec_imp; a.rec = h->rec;
    a.single_group = h->single_group;
    a.no_bond = h->no_bond;
    return a;
}

static int export_active(vg_scene *h, ActiveOut *out, size_t cap, unsigned long long *dcnt)
{
    HIPCHK(hipMemsetAsync(dcnt, 0, sizeof(unsigned long long), h->stream));
    hipLaunchKernelGGL(k_export_active, dim3(nblk(h->n, 256)), dim3(256), 0, h->stream, h->geo[h->cur],
                       h->ell[h->ecur], h->n, h->ki, (const unsigned *)h->ocount, (const unsigned *)h->onbr,
                       (const WallDev *)h->walls, (const unsigned *)h->d_wall_ids,
                       (const unsigned *)h->d_sphere_ids,
                       h->ent_valid ? (const double *)h->rec : (const double *)nullptr, out,
                       (unsigned long long)cap, dcnt);
    HIPCHK(hipGetLastError());
    return 0;
}

// RebuildIactRecordTab before the candidate-list reset (NBSConsultant.cpp:40-47):
// the active contacts and the wall impulses of the last step are kept, since
// the rebuild re-sorts the spheres and their lists.
static int snapshot_records(vg_scene *h)
{
    unsigned long long *dcnt = nullptr;
    if (dalloc(h, &dcnt, 1)) return VG_ENOMEM;
    if (export_active(h, nullptr, 0, dcnt)) return VG_EHIP;
    unsigned long long cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, dcnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (cnt > h->snap_cap) {
        if (h->snap) HIPCHK(hipFree(h->snap));
        h->snap_cap = cnt + cnt / 4 + 64;
        if (dalloc(h, &h->snap, h->snap_cap)) return VG_ENOMEM;
    }
    if (export_active(h, h->snap, h->snap_cap, dcnt)) return VG_EHIP;
    if (h->nw && h->ent_valid)
        hipLaunchKernelGGL(k_wall_impulse, dim3(h->nw), dim3(256), 0, h->stream, h->ell[h->ecur], h->n, h->ki,
                           (const double *)h->rec, h->snap_wimp);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(dcnt));
    h->snap_count = (long long)cnt;
    h->snap_valid = true;
    return 0;
}

static int one_step(vg_scene *h, bool rec, bool assign_spheres)
{
    StepArgs a = step_args(h);
    const bool kt = 2 * (h->kev_used + 1) <= (int)h->kev.size();
    if (kt) HIPCHK(hipEventRecord(h->kev[2 * h->kev_used], h->stream));
    const dim3 grid(nblk(h->n, 256)), blk(256);
    if (h->sd_only) {
        if (rec) hipLaunchKernelGGL((k_step<true, true>), grid, blk, 0, h->stream, a);
        else hipLaunchKernelGGL((k_step<false, true>), grid, blk, 0, h->stream, a);
    } else {
        if (rec) hipLaunchKernelGGL((k_step<true, false>), grid, blk, 0, h->stream, a);
        else hipLaunchKernelGGL((k_step<false, false>), grid, blk, 0, h->stream, a);
    }
    if (kt) {
        HIPCHK(hipEventRecord(h->kev[2 * h->kev_used + 1], h->stream));
        h->kev_used++;
    }
    h->last_launches++;
    h->cur ^= 1;
    h->step_count++;
    h->round++;
    h->t += h->dt;   // SystemParameter::NextStep
    if (h->nw) {
        hipLaunchKernelGGL(k_walls, dim3(nblk(h->nw, 64)), dim3(64), 0, h->stream, h->walls, h->nw,
                           h->assigned, h->walls_ref, h->wkin, h->dt, h->t, h->flags, a.step, a.relist_half);
    }
    if (assign_spheres) {
        // SimMediator::Run(map): the assigned statuses replace the integrated
        // ones before contact detection and the rebuild; positions may jump,
        // so every sphere relists at the next step
        hipLaunchKernelGGL(k_assign_spheres, dim3(nblk(h->n, 256)), dim3(256), 0, h->stream, h->geo[h->cur], h->n,
                           (const double *)h->assigned_sph, rec ? h->rec_imp : (double *)nullptr);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, h->stream, h->flags + (h->step_count % 3), 1ll, 1);
    }
    h->rec_valid = rec;
    h->ent_valid = rec;
    h->snap_valid = false;
    HIPCHK(hipGetLastError());
    if (h->round % h->up_iact == 0) {
        if (rec && snapshot_records(h)) return VG_EHIP;
        return rebuild(h);
    }
    return 0;
}

extern "C" int vg_step(vg_scene *h, int nsteps)
{
    if (!h || nsteps < 0) return VG_EINVAL;
    h->last_launches = 0;
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    for (int k = 0; k < nsteps; ++k) {
        int rc = one_step(h, k == nsteps - 1, false);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    h->timing = true;
    return 0;
}

extern "C" int vg_sync(vg_scene *h)
{
    if (!h) return VG_EINVAL;
    HIPCHK(hipStreamSynchronize(h->stream));
    int flags[8];
    HIPCHK(hipMemcpy(flags, h->flags, sizeof(flags), hipMemcpyDeviceToHost));
    if (flags[3] & ERR_INNER_CAP)
        return fail(h, VG_ECAPACITY, "neighbour list capacity exceeded (set VG_NEIGHBOUR_CAPACITY)");
    if (h->timing) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, h->ev0, h->ev1) == hipSuccess) h->last_ms = ms;
        h->timing = false;
    }
    return 0;
}

extern "C" int vg_step_assigned(vg_scene *h, int n, const unsigned *ids, const double *pos, const double *vel,
                                const double *ox, const double *oz, const double *av)
{
    if (!h || n < 0 || (n > 0 && (!ids || !pos || !vel || !ox || !oz || !av))) return VG_EINVAL;
    std::vector<double> buf((size_t)15 * h->nw, 0.0);
    std::vector<char> set(h->nw, 0);
    std::vector<double> sb;
    bool any_sphere = false;
    for (int k = 0; k < n; ++k) {
        if (ids[k] >= (unsigned)h->n_do)
            return fail(h, VG_EINVAL, "assigned status of an unknown DO id");
        if (h->do_to_sphere[ids[k]] >= 0) {
            if (!any_sphere) sb.assign((size_t)16 * h->n, 0.0);
            any_sphere = true;
            double *b = &sb[(size_t)16 * h->do_to_sphere[ids[k]]];
            for (int c = 0; c < 3; ++c) {
                b[c] = pos[3 * k + c]; b[3 + c] = vel[3 * k + c]; b[6 + c] = ox[3 * k + c];
                b[9 + c] = oz[3 * k + c]; b[12 + c] = av[3 * k + c];
            }
            b[15] = 1.0;
            continue;
        }
        const int w = h->do_to_wall[ids[k]];
        double *b = &buf[(size_t)15 * w];
        for (int c = 0; c < 3; ++c) {
            b[c] = pos[3 * k + c]; b[3 + c] = vel[3 * k + c]; b[6 + c] = ox[3 * k + c];
            b[9 + c] = oz[3 * k + c]; b[12 + c] = av[3 * k + c];
        }
        set[w] = 1;
    }
    if (any_sphere) {
        if (!h->assigned_sph && dalloc(h, &h->assigned_sph, (size_t)16 * h->n)) return VG_ENOMEM;
        HIPCHK(hipMemcpyAsync(h->assigned_sph, sb.data(), sizeof(double) * sb.size(), hipMemcpyHostToDevice,
                              h->stream));
    }
    bool changed = false;
    for (int w = 0; w < h->nw; ++w) {
        const int want = set[w] ? 1 : (h->walls_h[w].kin == 1 ? 0 : h->walls_h[w].kin);
        if (want != h->walls_h[w].kin) {
            h->walls_h[w].kin = want;
            changed = true;
        }
    }
    if (h->nw) {
        HIPCHK(hipMemcpyAsync(h->assigned, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice, h->stream));
        if (changed) {
            // only the kin field changes here; the device copy holds the live status
            std::vector<WallDev> cur(h->nw);
            HIPCHK(hipMemcpyAsync(cur.data(), h->walls, sizeof(WallDev) * h->nw, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            for (int w = 0; w < h->nw; ++w) cur[w].kin = h->walls_h[w].kin;
            HIPCHK(hipMemcpyAsync(h->walls, cur.data(), sizeof(WallDev) * h->nw, hipMemcpyHostToDevice, h->stream));
        }
    }
    h->last_launches = 0;
    int rc = one_step(h, true, any_sphere);
    if (rc) return rc;
    // synchronous: the host buffers are reused by the next call
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int vg_system_energy(vg_scene *h, double *out)
{
    if (!h || !out) return VG_EINVAL;
    const int n = h->n;
    const int nb = (int)std::max(1u, std::min(1024u, nblk(n, 256)));
    double *avg = h->en_out + 16;
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_energy_part, dim3(nb), dim3(256), 0, h->stream, h->geo[h->cur], n,
                           (const double2 *)h->mm, h->g[0], h->g[1], h->g[2], (const double *)avg, pass, h->en_part);
        hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, h->stream, (const double *)h->en_part, nb, n, pass,
                           avg, h->en_out);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, h->en_out, 16 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int vg_set_wall_rotation(vg_scene *h, int n, const unsigned *ids, const double *c, const double *a,
                                    double omega, double t_ref)
{
    if (!h || n < 0) return VG_EINVAL;
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<WallDev> cur(h->nw);
    std::vector<WallKin> kin(h->nw);
    if (h->nw) HIPCHK(hipMemcpy(cur.data(), h->walls, sizeof(WallDev) * h->nw, hipMemcpyDeviceToHost));
    const double al = olen(a);
    for (int k = 0; k < n; ++k) {
        if (ids[k] >= (unsigned)h->n_do || h->do_to_wall[ids[k]] < 0)
            return fail(h, VG_EINVAL, "wall rotation applies to walls only");
        const int w = h->do_to_wall[ids[k]];
        kin[w].c = mk(c[0], c[1], c[2]);
        kin[w].a = mk(a[0] / al, a[1] / al, a[2] / al);
        kin[w].omega = omega;
        kin[w].t_ref = t_ref;
        cur[w].kin = 2;
        h->walls_h[w].kin = 2;
    }
    if (h->nw) {
        HIPCHK(hipMemcpy(h->walls_ref, cur.data(), sizeof(WallDev) * h->nw, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->walls, cur.data(), sizeof(WallDev) * h->nw, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->wkin, kin.data(), sizeof(WallKin) * h->nw, hipMemcpyHostToDevice));
    }
    return 0;
}

extern "C" int vg_get_state(vg_scene *h, double *pos, double *vel, double *ox, double *oz, double *av,
                            double *impact, double *ang_impact)
{
    if (!h) return VG_EINVAL;
    HIPCHK(hipStreamSynchronize(h->stream));
    const int nd = h->n_do, n = h->n;
    double *dout = nullptr;
    if (dalloc(h, &dout, (size_t)21 * n)) return VG_ENOMEM;
    std::vector<double> sb((size_t)21 * n, 0.0);
    hipLaunchKernelGGL(k_scatter_state, dim3(nblk(n, 256)), dim3(256), 0, h->stream, h->geo[h->cur], n,
                       h->rec_valid ? h->rec_imp : (const double *)nullptr, dout);
    std::vector<double> wimp((size_t)6 * std::max(h->nw, 1), 0.0);
    double *dw = nullptr;
    if (h->nw && h->snap_valid) {
        if (dalloc(h, &dw, (size_t)6 * h->nw)) return VG_ENOMEM;
        HIPCHK(hipMemcpyAsync(dw, h->snap_wimp, sizeof(double) * 6 * h->nw, hipMemcpyDeviceToDevice, h->stream));
    } else if (h->nw && h->ent_valid) {
        if (dalloc(h, &dw, (size_t)6 * h->nw)) return VG_ENOMEM;
        hipLaunchKernelGGL(k_wall_impulse, dim3(h->nw), dim3(256), 0, h->stream, h->ell[h->ecur], n, h->ki,
                           (const double *)h->rec, dw);
    }
    HIPCHK(hipMemcpyAsync(sb.data(), dout, sizeof(double) * sb.size(), hipMemcpyDeviceToHost, h->stream));
    if (dw) HIPCHK(hipMemcpyAsync(wimp.data(), dw, sizeof(double) * 6 * h->nw, hipMemcpyDeviceToHost, h->stream));
    std::vector<WallDev> walls(h->nw);
    if (h->nw)
        HIPCHK(hipMemcpyAsync(walls.data(), h->walls, sizeof(WallDev) * h->nw, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(dout));
    if (dw) HIPCHK(hipFree(dw));
    std::vector<double> hb((size_t)21 * nd, 0.0);
    for (int s = 0; s < n; ++s)
        std::memcpy(&hb[(size_t)21 * h->sphere_ids[s]], &sb[(size_t)21 * s], 21 * sizeof(double));
    for (int w = 0; w < h->nw; ++w) {
        const unsigned id = h->wall_ids[w];
        double *d = &hb[(size_t)21 * id];
        const WallDev &W = walls[w];
        const d3 v[5] = {W.pos, W.vel, W.ox, W.oz, W.av};
        for (int k = 0; k < 5; ++k) { d[3 * k] = v[k].x; d[3 * k + 1] = v[k].y; d[3 * k + 2] = v[k].z; }
        // assigned statuses carry zero stored impulses (DOContainer::SetDOStatus)
        const bool zero = W.kin != 0;
        for (int c = 0; c < 6; ++c) d[15 + c] = zero ? 0.0 : wimp[(size_t)6 * w + c];
    }
    double *outs[7] = {pos, vel, ox, oz, av, impact, ang_impact};
    for (int f = 0; f < 7; ++f) {
        if (!outs[f]) continue;
        for (int i = 0; i < nd; ++i)
            for (int c = 0; c < 3; ++c) outs[f][3 * i + c] = hb[(size_t)21 * i + 3 * f + c];
    }
    return 0;
}

extern "C" double vg_time(const vg_scene *h) { return h ? h->t : 0.0; }

extern "C" int vg_num_interactions(vg_scene *h, long long *n)
{
    if (!h || !n) return VG_EINVAL;
    if (h->n_iact < 0) {
        unsigned long long *d = nullptr, c[2] = {0, 0};
        if (dalloc(h, &d, 2)) return VG_ENOMEM;
        HIPCHK(hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), h->stream));
        hipLaunchKernelGGL(k_count_iact, dim3(nblk(h->n, 256)), dim3(256), 0, h->stream, (const unsigned *)h->ocount,
                           (const unsigned *)h->onbr, h->n, h->ko, d);
        HIPCHK(hipMemcpyAsync(c, d, sizeof(c), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipFree(d));
        h->n_iact = (long long)(c[0] / 2 + c[1]);   // sphere pairs appear in both spheres' lists
    }
    *n = h->n_iact;
    return 0;
}

extern "C" int vg_set_skin(vg_scene *h, double skin)
{
    if (!h || !(skin >= 0.0)) return VG_EINVAL;
    HIPCHK(hipStreamSynchronize(h->stream));
    h->skin = std::max(skin, h->bond_reach * 1.01);
    // every sphere rebuilds its neighbour list with the new skin at the next step
    int one = 1;
    HIPCHK(hipMemcpy(h->flags + (h->step_count % 3), &one, sizeof(int), hipMemcpyHostToDevice));
    return 0;
}

extern "C" double vg_get_skin(const vg_scene *h) { return h ? h->skin : 0.0; }

extern "C" int vg_get_active(vg_scene *h, long long *n, unsigned *lo, unsigned *hi, vg_status *st)
{
    if (!h || !n) return VG_EINVAL;
    HIPCHK(hipStreamSynchronize(h->stream));
    unsigned long long cnt = 0;
    std::vector<ActiveOut> ho;
    if (h->snap_valid) {
        cnt = (unsigned long long)h->snap_count;
        if (st) {
            ho.resize(std::min<size_t>(cnt, (size_t)std::max(*n, 0ll)));
            if (!ho.empty()) HIPCHK(hipMemcpy(ho.data(), h->snap, sizeof(ActiveOut) * ho.size(), hipMemcpyDeviceToHost));
        }
    } else {
        unsigned long long *dcnt = nullptr;
        if (dalloc(h, &dcnt, 1)) return VG_ENOMEM;
        if (export_active(h, nullptr, 0, dcnt)) return VG_EHIP;
        HIPCHK(hipMemcpyAsync(&cnt, dcnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (st) {
            const size_t cap = (size_t)std::max(*n, 0ll);
            ActiveOut *dout = nullptr;
            if (dalloc(h, &dout, cap)) return VG_ENOMEM;
            if (export_active(h, dout, cap, dcnt)) return VG_EHIP;
            ho.resize(std::min<size_t>(cnt, cap));
            if (!ho.empty()) HIPCHK(hipMemcpyAsync(ho.data(), dout, sizeof(ActiveOut) * ho.size(), hipMemcpyDeviceToHost,
                                                  h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            HIPCHK(hipFree(dout));
        }
        HIPCHK(hipFree(dcnt));
    }
    if (!st) {
        *n = (long long)cnt;
        return 0;
    }
    if ((long long)cnt > *n) return fail(h, VG_ECAPACITY, "vg_get_active: capacity smaller than the active count");
    std::sort(ho.begin(), ho.end(), [](const ActiveOut &x, const ActiveOut &y) {
        return x.lo != y.lo ? x.lo < y.lo : x.hi < y.hi;
    });
    for (size_t k = 0; k < ho.size(); ++k) {
        const ActiveOut &A = ho[k];
        lo[k] = A.lo;
        hi[k] = A.hi;
        vg_status &S = st[k];
        S.contact = (A.flags & F_CONTACT) != 0;
        S.last_contact = (A.flags & F_LASTCONTACT) != 0;
        S.bond = (A.flags & F_BOND) != 0;
        S.last_bond = (A.flags & F_LASTBOND) != 0;
        S.shear[0] = A.hist[0]; S.shear[1] = A.hist[1]; S.shear[2] = A.hist[2];
        // the fourth history word is Kn for ISwH and the initial velocity for ISwHertz
        S.kn = A.hist[3];
        S.init_vel = A.hist[3];
        for (int c = 0; c < 3; ++c) {
            S.point[c] = A.rec[c]; S.dir[c] = A.rec[3 + c]; S.to_master[c] = A.rec[6 + c];
            S.ang_to_master[c] = A.rec[9 + c];
        }
        S.overlap = A.rec[12];
    }
    *n = (long long)ho.size();
    return 0;
}

extern "C" int vg_set_history(vg_scene *h, int n, const unsigned *lo, const unsigned *hi, const vg_status *st)
{
    (void)lo; (void)hi; (void)st;
    if (!h) return VG_EINVAL;
    if (n == 0) return 0;
    return fail(h, VG_EUNSUPPORTED, "history restore is not implemented yet");
}

extern "C" int vg_step_timing(vg_scene *h, double *ms, int *launches)
{
    if (!h) return VG_EINVAL;
    if (h->kev_used > 0) {
        HIPCHK(hipStreamSynchronize(h->stream));
        double tot = 0.0;
        for (int k = 0; k < h->kev_used; ++k) {
            float m = 0.0f;
            HIPCHK(hipEventElapsedTime(&m, h->kev[2 * k], h->kev[2 * k + 1]));
            tot += m;
        }
        if (ms) *ms = tot;
        if (launches) *launches = h->kev_used;
        return 0;
    }
    if (ms) *ms = h->last_ms;
    if (launches) *launches = h->last_launches;
    return 0;
}

extern "C" int vg_kernel_timing(vg_scene *h, int max_launches)
{
    if (!h || max_launches < 0) return VG_EINVAL;
    HIPCHK(hipStreamSynchronize(h->stream));
    for (auto e : h->kev) HIPCHK(hipEventDestroy(e));
    h->kev.clear();
    h->kev_used = 0;
    for (int k = 0; k < 2 * max_launches; ++k) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        h->kev.push_back(e);
    }
    return 0;
}

extern "C" int vg_stats(vg_scene *h, long long *rebuilds, long long *relists, int *max_fill, int *capacity)
{
    if (!h) return VG_EINVAL;
    HIPCHK(hipStreamSynchronize(h->stream));
    unsigned long long r = 0;
    int flags[8];
    HIPCHK(hipMemcpy(&r, h->relists, sizeof(r), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(flags, h->flags, sizeof(flags), hipMemcpyDeviceToHost));
    if (rebuilds) *rebuilds = h->rebuilds;
    if (relists) *relists = (long long)r;
    if (max_fill) *max_fill = flags[4];
    if (capacity) *capacity = h->ki;
    return 0;
}

// Streaming copy of `bytes` bytes, `iters` times, on the current device: the
// achievable HBM bandwidth (read + write bytes / time) and the calibration
// kernel for HBM traffic counters.  No reference counterpart (diagnostics).
extern "C" int vg_copy_bench(size_t bytes, int iters, double *gbps)
{
    if (bytes < 1024 || iters < 1 || !gbps) return VG_EINVAL;
    const size_t n = bytes / sizeof(double2);
    double2 *s = nullptr, *d = nullptr;
    if (hipMalloc((void **)&s, n * sizeof(double2)) != hipSuccess) return VG_ENOMEM;
    if (hipMalloc((void **)&d, n * sizeof(double2)) != hipSuccess) {
        (void)hipFree(s);
        return VG_ENOMEM;
    }
    (void)hipMemset(s, 0, n * sizeof(double2));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const unsigned blocks = 256 * 8 * 4;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, s, d, n);   // warm-up
    (void)hipEventRecord(e0, 0);
    for (int k = 0; k < iters; ++k) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, s, d, n);
    (void)hipEventRecord(e1, 0);
    const hipError_t rc = hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *gbps = 2.0 * (double)n * sizeof(double2) * iters / (ms * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(s);
    (void)hipFree(d);
    return rc == hipSuccess ? 0 : VG_EHIP;
}
