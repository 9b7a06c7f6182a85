// vg_api.hip -- host side of libvedo_hip.so: the C-ABI of include/vedo_gpu.h.
//
// Plays the role of VEDO's SimMediator (SimMediator.cpp:640-842) and
// LeapConsultant (LeapConsultant.cpp:49-195) on one GPU for a batch of
// independent scenes: holds the device state of every scene in one packed
// layout, orders the per-step kernels on its own stream (one step kernel for
// all scenes), and runs the candidate-list rebuild every up_iact steps.
// vg_create makes a batch of one scene; vg_batch_create / vg_admit make the
// batched and served forms.  A scene's arithmetic does not depend on the
// other scenes of its batch: its grid, IactModel table, walls, time step and
// rebuild cadence are its own, and its per-sphere sums follow its own
// candidate order, so its result is the same alone and in a batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include <rocprim/rocprim.hpp>
#include <string>
#include <vector>

#include "../../include/vedo_gpu.h"
#include "vg_kernels.h"
// one translation unit: the kernels are launched from this file
#include "vg_kernels.hip"


using namespace vg;

static std::string g_create_error;

// Host record of one scene slot.
struct SceneHost {
    int state = VG_SCENE_FREE;
    // DO bookkeeping (DO ids are the scene's own)
    int n_do = 0, n = 0, nw = 0;
    std::vector<unsigned> sphere_ids;   // rank -> DO id
    std::vector<unsigned> wall_ids;     // wall order -> DO id
    std::vector<int> do_to_sphere;      // DO id -> rank or -1
    std::vector<int> do_to_wall;        // DO id -> wall slot (batch-wide) or -1
    std::vector<int> do_model;          // DO id -> DOModel index (the description's)
    std::vector<char> model_is_sphere;  // per DOModel of the description
    bool all_sphere_laws = true;        // an IactModel for every pair of sphere groups
    int w_begin = 0, law_off = 0, n_groups = 0, model_off = 0, n_models = 0;
    double dt = 0, t = 0, t_stop = 0, g[3] = {0, 0, 0}, sf = 1.1, skin = 0, bond_reach = 0, rmin = 0;
    bool sd_only = true, single_group = true, no_bond = true;
    int group0 = 0;                     // the spheres' group where they share one
    bool h_only = true;                 // every law present is ISwH
    bool has_constrained = false;       // a wall moves under its contacts (DOConstrained*)
    int zon = 0;                        // zone of interest: bounded axes (bits), bounds
    double zlo[3] = {0, 0, 0}, zhi[3] = {0, 0, 0};
    long long steps_done = 0, steps_target = -1;
    int seg_begin = 0, seg_end = 0;     // slots of the spheres in the current layout
    // initial state until the scene is placed
    std::vector<double4> pos0, vel0, avl0;
    std::vector<double> ori0;
    std::vector<int> mdl0;
    std::vector<WallDev> walls0;
    std::vector<double4> holes0;        // per wall: hole radius, x / y offset, 1 for a holed plate
    bool has_hole = false;              // a QuasiPlateWithCircularHole among the walls
    bool has_pbc = false;               // periodic boundary (PeriodicBoundaryConditions)
    PbcDev pbc0{};
    bool nbs = false;                   // NBSConsultant's candidate list (else LeapConsultant's)
    double nbs_range = 0.0;             // the range NBS takes for the zone size: its last local DO's
    std::vector<WallKin> kin0;          // rotations prescribed before placement
    std::vector<char> kin0_set;
    // result kept on the host once a finished scene leaves the device
    std::vector<double> result;         // [n_do][21]
    bool failed = false;                // blew up: left the batch without a result
};

// A captured rebuild period of a batch handle (graph_period).
struct StepGraph {
    std::vector<unsigned char> key;   // launch arguments it was captured with
    hipGraphExec_t exec = nullptr;
    unsigned long long last_use = 0;
};

static void graph_free(StepGraph &g)
{
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
    g.key.clear();
}

struct vg_scene {
    std::string err;
    hipStream_t stream = nullptr;
    int device = 0;
    bool use_graphs = true;             // batch handles replay captured rebuild periods (vg_set_graphs)
    std::vector<StepGraph> graphs = std::vector<StepGraph>(8);
    unsigned long long graph_clock = 0;
    long long graph_replays = 0;
    bool batch = false;                 // made by vg_batch_create
    // capacities and sizes
    int ld = 0, wcap = 0, scap = 0, ki = 48;
    int n = 0;                          // occupied sphere slots (placed scenes)
    int nw_used = 0;                    // wall slots up to the last one in use
    std::vector<SceneHost> sc;          // scene slots
    std::vector<SceneDev> sdev;         // host copy of the device scene table
    std::vector<ModelDev> models_h;
    std::vector<double2> mm_h;          // per model: mass, mass moment of inertia (spheres)
    std::vector<LawDev> laws_h;
    std::vector<WallDev> walls_h;       // host copy of wall kinematic modes (kin)
    std::vector<char> wall_used;
    // batch-wide properties of the placed scenes
    bool sd_only = true, h_only = true;
    int single_group = 1, no_bond = 1;
    int precision = VG_PRECISION_FP64;   // vg_set_precision
    bool skin_auto = false;             // vg_set_skin_auto: skins follow each scene's speed
    double skin_floor = 0.0, skin_cap = 0.0;
    unsigned up_iact = 1;
    unsigned long long round = 0, step_count = 0;
    long long rebuilds = 0;
    long long n_iact = 0;
    bool record = true;                 // per-entry ImpactStatus records (one-scene handles)
    // device
    Geo geo[2];
    unsigned *oid_spare = nullptr;      // rank / model arrays the next rebuild permutes into
    int *mdl_spare = nullptr;
    int cur = 0;
    Ell ell[2];
    int ecur = 0;
    ModelDev *models = nullptr;
    double2 *mm = nullptr;
    int models_cap = 0;
    LawDev *laws = nullptr;
    int laws_cap = 0;
    SceneDev *scenes = nullptr;
    double *stime = nullptr;            // [scap][2] scene times for the wall kernel
    int *rflags = nullptr;              // [scap][3] relist rings
    WallDev *walls = nullptr, *walls_ref = nullptr;
    WallKin *wkin = nullptr;
    double4 *holes = nullptr;           // [wcap] hole table; allocated when a scene with a holed plate is placed
    PbcDev *pbc = nullptr;              // periodic boundary of the scene; allocated when a periodic scene is placed
    bool nbs = false;                   // the scene takes NBSConsultant's candidate list
    PbcDev pbc_h{};                     // its host copy (traversal axes updated at every rebuild)
    // candidate list under a periodic boundary: occupied cells, pairs, list entries
    unsigned long long *pbc_ckey[2] = {nullptr, nullptr};
    unsigned *pbc_cfirst[2] = {nullptr, nullptr};
    unsigned *pbc_ncell = nullptr, *pbc_pcount = nullptr, *pbc_poff = nullptr, *pbc_first = nullptr;
    uint2 *pbc_pairs = nullptr;
    unsigned *pbc_ekey[2] = {nullptr, nullptr}, *pbc_eval[2] = {nullptr, nullptr};
    size_t pbc_cells_cap = 0, pbc_pairs_cap = 0;
    double *assigned = nullptr;
    double *assigned_sph = nullptr;     // [ld][16] assigned sphere statuses by rank (allocated on first use)
    double *en_part = nullptr, *en_out = nullptr;
    double *fetch = nullptr;            // state download scratch (grows)
    size_t fetch_cap = 0;
    double *pin = nullptr;              // pinned host staging of state downloads (grows)
    size_t pin_cap = 0;
    int *flags = nullptr;
    unsigned long long *relists = nullptr;
    unsigned long long *contacts = nullptr;   // per-wavefront contact-pair counts, summed by vg_stats
    unsigned long long *red = nullptr;  // [scap][8] bounds
    unsigned *inv = nullptr, *perm = nullptr, *svals = nullptr;
    unsigned long long *keys = nullptr, *skeys = nullptr;
    double *xs[2] = {nullptr, nullptr}, *ys[2] = {nullptr, nullptr}, *zs[2] = {nullptr, nullptr};
    long long *qidx = nullptr, *segs = nullptr;
    int *ovf = nullptr;                 // [scap] candidate-capacity overflow per scene
    unsigned *sid[2] = {nullptr, nullptr};   // scene slot of each sphere (quantile sorts)
    double *qval = nullptr;
    QselTarget *qsel_tg = nullptr;   // quantile selection: 6 targets per scene slot
    unsigned *qsel_hist = nullptr;   // 256 counters per target
    int *qmap = nullptr;             // scene slot -> live-scene index, -1

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
    // IactRecordTab snapshot taken before a rebuild re-sorts the spheres (one-scene handles)
    ActiveOut *snap = nullptr;
    size_t snap_cap = 0;
    long long snap_count = 0;
    double *snap_wimp = nullptr;
    bool snap_valid = false;
    // Constrained walls: each one's candidate pairs in IactPairTab order,
    // laid down at every rebuild; a step writes each pair's master impulse
    // with the step's stamp, and k_cw_sum adds a wall's pairs in that order.
    bool any_constrained = false;       // a placed scene has a constrained wall
    long long cw_cap = 0, cw_n = 0;     // pair slots allocated / in use
    double *cw_val = nullptr;           // [cw_cap][3]
    unsigned *cw_stamp = nullptr;       // [cw_cap]
    unsigned long long *cw_key[2] = {nullptr, nullptr};
    unsigned *cw_vidx[2] = {nullptr, nullptr};
    int *cw_w = nullptr, *cw_p = nullptr;    // [CW_SLOTS][ld] a sphere's constrained walls, their pair slots
    unsigned *cw_cnt = nullptr;         // [ld + 1] per-sphere pair counts, then offsets
    int *cw_beg = nullptr, *cw_end = nullptr;   // [wcap] pair range of each wall slot
    double *cw_imp = nullptr;           // [wcap][3] each constrained wall's impulse of the step
    // zone of interest: per-sphere erase flags, per-scene counts, ranks
    bool any_zoi = false;
    unsigned char *zoi_dead = nullptr;   // [ld]
    int *zoi_cnt = nullptr;             // [scap]
    unsigned *zoi_rank = nullptr, *zoi_pre = nullptr;   // [ld + 1] erase flags by rank, their prefix count
    // timing
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;
    int last_launches = 0;
    bool timing = false;
    // per-launch events around the step kernel (vg_kernel_timing)
    std::vector<hipEvent_t> kev;
    int kev_used = 0;
    std::vector<hipEvent_t> rev;   // candidate-list rebuilds bracketed while kernel timing is on
    int rev_used = 0;
    // the rebuild due after the last step of a vg_step call waits for the
    // handle's next use, so that the step's records stay readable in place
    bool rebuild_pending = false;
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

static Tables tables(vg_scene *h) { return Tables{h->models, h->laws}; }

static double model_range(const vg_model &m)
{
    // DOModel::Analysis (DOModel.cpp:17-149), global walls
    if (m.shape == VG_SPHERE) return m.attr[0];
    if (m.shape == VG_QUASICYLINDER) return m.scope == VG_LOCAL ? std::max(m.attr[0], 0.5 * m.attr[1]) : m.attr[0];
    if (m.shape == VG_QUASIPLATE || m.shape == VG_QUASIPLATE_HOLE) {   // DOModel.cpp:90-106
        if (m.scope == VG_LOCAL) return 0.5 * std::max(std::max(m.attr[0], m.attr[1]), m.attr[2]);
        return 0.5 * m.attr[1];
    }
    return 0.0;
}

static const double PI = 4.0 * std::atan(1.0);

#include <chrono>
// VG_DEBUG phase timing of the rebuild (host clock)
struct PhaseClock {
    bool on = getenv("VG_DEBUG") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void mark(const char *what)
    {
        if (!on) return;
        const auto now = std::chrono::steady_clock::now();
        fprintf(stderr, "[vg] %-28s %8.3f ms\n", what, std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
};

// ---------------------------------------------------------------------------
// Device tables
// ---------------------------------------------------------------------------

static int upload_scenes(vg_scene *h)
{
    HIPCHK(hipMemcpyAsync(h->scenes, h->sdev.data(), sizeof(SceneDev) * h->scap, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));   // the host table is reused right away
    return 0;
}

template <class T> static int grow_table(vg_scene *h, T **d, int *cap, const std::vector<T> &host)
{
    if ((int)host.size() > *cap) {
        HIPCHK(hipStreamSynchronize(h->stream));
        if (*d) HIPCHK(hipFree(*d));
        *d = nullptr;
        *cap = (int)host.size() * 2 + 16;
        if (dalloc(h, d, (size_t)*cap)) return VG_ENOMEM;
    }
    if (!host.empty()) HIPCHK(hipMemcpy(*d, host.data(), sizeof(T) * host.size(), hipMemcpyHostToDevice));
    return 0;
}

// batch-wide kernel specialisations follow the placed scenes
static void update_batch_flags(vg_scene *h)
{
    h->sd_only = true;
    h->h_only = true;
    h->single_group = 1;
    h->no_bond = 1;
    for (const SceneHost &S : h->sc) {
        if (S.state != VG_SCENE_RUNNING && S.state != VG_SCENE_FINISHED) continue;
        h->sd_only = h->sd_only && S.sd_only;
        h->h_only = h->h_only && S.h_only;
        h->single_group = h->single_group && S.single_group;
        h->no_bond = h->no_bond && S.no_bond;
    }
    h->any_constrained = false;
    h->any_zoi = false;
    for (const SceneHost &S : h->sc)
        if (S.state == VG_SCENE_RUNNING || S.state == VG_SCENE_FINISHED || S.state == VG_SCENE_QUEUED) {
            h->any_constrained = h->any_constrained || S.has_constrained;
            h->any_zoi = h->any_zoi || S.zon != 0;
        }
}

// ---------------------------------------------------------------------------
// Scene description -> host scene (validation as the reference's setup)
// ---------------------------------------------------------------------------

static int scene_from_desc(vg_scene *h, const vg_scene_desc *d, SceneHost &S, std::string &err)
{
    if (d->consultant != 0 && d->consultant != 1) {
        err = "consultant: 0 (LeapConsultant) or 1 (NBSConsultant)";
        return VG_EINVAL;
    }
    if (d->consultant == 1 && h->batch) {
        err = "NBSConsultant's candidate list: only in a handle of one scene (vg_create)";
        return VG_EUNSUPPORTED;
    }
    if (d->n_do <= 0 || d->n_models <= 0 || d->n_groups <= 0 || !d->models || !d->iact || !d->model_of ||
        !d->pos || !d->vel || !d->ox || !d->oz || !d->av) {
        err = "incomplete scene description";
        return VG_EINVAL;
    }
    S = SceneHost();
    S.nbs = d->consultant == 1;
    S.n_do = d->n_do;
    S.n_groups = d->n_groups;
    S.n_models = d->n_models;
    S.do_to_sphere.assign(d->n_do, -1);
    S.do_to_wall.assign(d->n_do, -1);
    S.do_model.assign(d->model_of, d->model_of + d->n_do);
    for (int i = 0; i < d->n_do; ++i) {
        const int m = d->model_of[i];
        if (m < 0 || m >= d->n_models) {
            err = "DOStatus names an unknown DOModel";
            return VG_EINVAL;
        }
        const vg_model &M = d->models[m];
        if (M.density < 0.0) {
            err = "negative density";
            return VG_EUNSUPPORTED;
        }
        if (M.group < 0 || M.group >= d->n_groups) {
            err = "DOModel names an unknown DOGroup";
            return VG_EINVAL;
        }
        if (M.shape == VG_SPHERE) {
            if (M.scope != VG_LOCAL || M.behavior != VG_MOBILE || M.density == 0.0) {
                err = "spheres must be local, mobile and of positive density";
                return VG_EUNSUPPORTED;
            }
            S.do_to_sphere[i] = (int)S.sphere_ids.size();
            S.sphere_ids.push_back((unsigned)i);
        } else if (M.shape == VG_QUASIPLATE || M.shape == VG_QUASICYLINDER || M.shape == VG_QUASIPLATE_HOLE) {
            if (M.scope != VG_GLOBAL ||
                (M.behavior != VG_FIXED && M.behavior != VG_ORBITAL && M.behavior != VG_CONSTRAINED)) {
                err = "walls must be global and fixed, orbital or constrained";
                return VG_EUNSUPPORTED;
            }
            if (M.behavior == VG_CONSTRAINED && !(M.density * M.density_factor > 0.0)) {
                err = "a constrained wall needs a positive mass (density and density factor)";
                return VG_EUNSUPPORTED;
            }
            S.do_to_wall[i] = (int)S.wall_ids.size();   // wall order; the slot offset is added at admission
            S.wall_ids.push_back((unsigned)i);
        } else {
            err = "shape is not on the GPU path";
            return VG_EUNSUPPORTED;
        }
    }
    S.n = (int)S.sphere_ids.size();
    S.nw = (int)S.wall_ids.size();
    if (S.n == 0) {
        err = "scene has no spheres";
        return VG_EINVAL;
    }
    double bond_reach = 0.0;
    for (int k = 0; k < d->n_groups * d->n_groups; ++k) {
        const int L = d->iact[k].law;
        if (L < 0 || L > 5) {
            err = "unknown impact solver";
            return VG_EINVAL;
        }
        if ((L == VG_BSDBF || L == VG_BSDB2F) && d->iact[k].p[4] != 0.0 && d->iact[k].p[0] != 0.0)
            bond_reach = std::max(bond_reach, std::fabs(d->iact[k].p[4] / d->iact[k].p[0]));
        if (L == VG_H || L == VG_HERTZ) S.sd_only = false;
        if (L != 0 && L != VG_H) S.h_only = false;
        if ((L == VG_BSDBF || L == VG_BSDB2F || L == VG_BTSDBF) && d->iact[k].p[4] != 0.0) S.no_bond = false;
    }
    {
        // which groups hold spheres, and does every pair of them interact
        std::vector<char> sg((size_t)d->n_groups, 0);
        S.model_is_sphere.assign((size_t)d->n_models, 0);
        for (int m = 0; m < d->n_models; ++m)
            if (d->models[m].shape == VG_SPHERE) {
                S.model_is_sphere[(size_t)m] = 1;
                sg[(size_t)d->models[m].group] = 1;
            }
        for (int a = 0; a < d->n_groups; ++a)
            for (int b = 0; b < d->n_groups; ++b)
                if (sg[(size_t)a] && sg[(size_t)b] && d->iact[a * d->n_groups + b].law == 0) S.all_sphere_laws = false;
    }
    {
        int g0 = -1;
        for (unsigned id : S.sphere_ids) {
            const int g = d->models[d->model_of[id]].group;
            if (g0 < 0) g0 = g;
            else if (g != g0) S.single_group = false;
        }
        S.group0 = g0 < 0 ? 0 : g0;
    }
    S.dt = d->dt; S.t = d->t_current; S.t_stop = d->t_stop;
    for (int c = 0; c < 3; ++c) S.g[c] = d->g[c];
    S.sf = d->safety_factor;
    // Boundary::Correct (Boundary.cpp:79-129): bounds sorted per axis, an axis bounded where they differ
    S.zon = 0;
    for (int c = 0; c < 3; ++c) {
        S.zlo[c] = std::min(d->zoi_lo[c], d->zoi_hi[c]);
        S.zhi[c] = std::max(d->zoi_lo[c], d->zoi_hi[c]);
        if (S.zlo[c] != S.zhi[c]) S.zon |= 1 << c;
    }
    // the periodic boundary likewise (SystemParameter PeriodicBoundaryConditions)
    S.pbc0 = PbcDev{};
    for (int c = 0; c < 3; ++c) {
        S.pbc0.lo[c] = std::min(d->pbc_lo[c], d->pbc_hi[c]);
        S.pbc0.hi[c] = std::max(d->pbc_lo[c], d->pbc_hi[c]);
        S.pbc0.len[c] = S.pbc0.hi[c] - S.pbc0.lo[c];
        S.pbc0.on[c] = S.pbc0.lo[c] != S.pbc0.hi[c];
    }
    S.has_pbc = S.pbc0.on[0] || S.pbc0.on[1] || S.pbc0.on[2];
    if (S.nbs && S.zon) {
        // NBSConsultant never erases DOs here: the zone of interest is a
        // LeapConsultant path (its NextStep's CleanUp, LeapConsultant.cpp:58-68)
        err = "a zone of interest with NBSConsultant's candidate list is not supported";
        return VG_EUNSUPPORTED;
    }
    if (S.nbs && !S.sphere_ids.empty()) {
        // NBSConsultant::Reset keeps the range of the last local DO it meets
        // as the zone size's radius (NBSConsultant.cpp:136-141)
        const vg_model &M = d->models[d->model_of[S.sphere_ids.back()]];
        S.nbs_range = model_range(M);
    }
    if (S.has_pbc) {
        if (h->batch) {
            err = "periodic boundaries: only in a handle of one scene (vg_create)";
            return VG_EUNSUPPORTED;
        }
        // EnforcePeriodicBoundaryConditions moves every DO (SimMediator.cpp:746): the
        // walls of a periodic scene are fixed and inside the range, where it leaves them
        for (int w = 0; w < S.nw; ++w) {
            const unsigned id = S.wall_ids[w];
            const vg_model &M = d->models[d->model_of[id]];
            bool inside = M.behavior == VG_FIXED;
            for (int c = 0; c < 3; ++c)
                if (S.pbc0.on[c])
                    inside = inside && d->pos[3 * id + c] >= S.pbc0.lo[c] && d->pos[3 * id + c] < S.pbc0.hi[c];
            if (!inside) {
                err = "periodic boundaries: walls must be fixed, with centres inside the periodic range";
                return VG_EUNSUPPORTED;
            }
        }
    }
    double rmin = 1e300;
    for (unsigned id : S.sphere_ids) rmin = std::min(rmin, d->models[d->model_of[id]].attr[0]);
    S.rmin = rmin;
    S.skin = d->inner_skin > 0 ? d->inner_skin : 0.1 * rmin;
    S.skin = std::max(S.skin, bond_reach * 1.01);
    S.bond_reach = bond_reach;
    // initial state (model indices are made batch-wide at admission)
    const int n = S.n;
    S.pos0.resize(n); S.vel0.resize(n); S.avl0.resize(n); S.ori0.resize(6 * (size_t)n); S.mdl0.resize(n);
    for (int s = 0; s < n; ++s) {
        const unsigned id = S.sphere_ids[s];
        const vg_model &M = d->models[d->model_of[id]];
        const double r = M.attr[0];
        const double mass = M.density * ((4.0 / 3.0 * PI) * r * r * r);   // DOModel.cpp:22-27
        const double sudo = M.density_factor * mass;
        S.pos0[s] = make_double4(d->pos[3 * id], d->pos[3 * id + 1], d->pos[3 * id + 2], r);
        S.vel0[s] = make_double4(d->vel[3 * id], d->vel[3 * id + 1], d->vel[3 * id + 2], 0.0);
        S.avl0[s] = make_double4(d->av[3 * id], d->av[3 * id + 1], d->av[3 * id + 2], sudo);
        for (int c = 0; c < 3; ++c) {
            S.ori0[(size_t)c * n + s] = d->ox[3 * id + c];
            S.ori0[(size_t)(3 + c) * n + s] = d->oz[3 * id + c];
        }
        S.mdl0[s] = d->model_of[id];
    }
    S.walls0.resize(S.nw);
    S.holes0.assign(S.nw, make_double4(0.0, 0.0, 0.0, 0.0));
    for (int w = 0; w < S.nw; ++w) {
        const unsigned id = S.wall_ids[w];
        const vg_model &M = d->models[d->model_of[id]];
        WallDev &W = S.walls0[w];
        W = WallDev();
        W.pos = mk(d->pos[3 * id], d->pos[3 * id + 1], d->pos[3 * id + 2]);
        W.vel = mk(d->vel[3 * id], d->vel[3 * id + 1], d->vel[3 * id + 2]);
        W.ox = mk(d->ox[3 * id], d->ox[3 * id + 1], d->ox[3 * id + 2]);
        W.oz = mk(d->oz[3 * id], d->oz[3 * id + 1], d->oz[3 * id + 2]);
        W.av = mk(d->av[3 * id], d->av[3 * id + 1], d->av[3 * id + 2]);
        W.a0 = M.attr[0]; W.a1 = M.attr[1]; W.a2 = M.attr[2];
        W.range = model_range(M);
        // a holed plate moves, bounds its contacts and enters the candidate
        // list as its plate (DOMap::CalDistance ignores the hole, DOMap.cpp:
        // 336-374); only its contact detection reads the hole table
        const bool holed = M.shape == VG_QUASIPLATE_HOLE;
        const bool plate = M.shape == VG_QUASIPLATE || holed;
        if (holed) {
            S.holes0[w] = make_double4(M.attr[3], M.attr[4], M.attr[5], 1.0);
            S.has_hole = true;
        }
        W.reach = plate ? 0.5 * std::sqrt(M.attr[0] * M.attr[0] + M.attr[1] * M.attr[1] + M.attr[2] * M.attr[2])
                        : std::sqrt(M.attr[0] * M.attr[0] + 0.25 * M.attr[1] * M.attr[1]);
        W.path = 0.0;
        W.shape = plate ? VG_QUASIPLATE : M.shape; W.group = M.group; W.behavior = M.behavior; W.kin = 0;
        // DOModel::Analysis (DOModel.cpp:56-84): the constrained walls' sudo-mass;
        // fixed and orbital walls carry none (DiscreteObject of DOFixed*/DOOrbital*)
        W.constrained = M.behavior == VG_CONSTRAINED;
        if (W.constrained) {
            double vol = plate ? M.attr[0] * M.attr[1] * M.attr[2] : PI * M.attr[0] * M.attr[0] * M.attr[1];
            // DOModel.cpp:108-116: the hole is cut out when it lies inside the plate
            if (holed && std::fabs(M.attr[4]) <= (0.5 * M.attr[0] - M.attr[3]) &&
                std::fabs(M.attr[5]) <= (0.5 * M.attr[2] - M.attr[3]))
                vol = M.attr[0] * M.attr[1] * M.attr[2] - M.attr[3] * M.attr[3] * PI * M.attr[1];
            W.msudo = M.density_factor * (M.density * vol);
        }
        W.fext[0] = M.fext[0]; W.fext[1] = M.fext[1]; W.fext[2] = M.fext[2];
        S.has_constrained = S.has_constrained || W.constrained;
    }
    (void)h;
    return 0;
}

// Reserve table entries, wall slots and a scene slot; the scene is placed on
// the device at the next rebuild boundary.
static int admit(vg_scene *h, const vg_scene_desc *d, long long steps, int *slot_out)
{
    SceneHost S;
    std::string err;
    int rc = scene_from_desc(h, d, S, err);
    if (rc) return fail(h, rc, err);
    if (d->up_iact != h->up_iact && !(d->up_iact == 0 && h->up_iact == 1))
        return fail(h, VG_EUNSUPPORTED, "the scenes of a batch share one UpIact");
    int slot = -1;
    for (int k = 0; k < h->scap; ++k)
        if (h->sc[k].state == VG_SCENE_FREE) {
            slot = k;
            break;
        }
    if (slot < 0) return fail(h, VG_ECAPACITY, "no free scene slot");
    long long spheres = 0;
    for (const SceneHost &T : h->sc)
        if (T.state == VG_SCENE_QUEUED || T.state == VG_SCENE_RUNNING || T.state == VG_SCENE_FINISHED)
            spheres += T.n;
    if (spheres + S.n > h->ld) return fail(h, VG_ECAPACITY, "sphere capacity of the batch exceeded");
    // a contiguous run of free wall slots (first fit)
    int w_begin = -1;
    for (int w = 0, run = 0; w < h->wcap && S.nw > 0; ++w) {
        run = h->wall_used[w] ? 0 : run + 1;
        if (run == S.nw) {
            w_begin = w - S.nw + 1;
            break;
        }
    }
    if (S.nw > 0 && w_begin < 0) return fail(h, VG_ECAPACITY, "wall capacity of the batch exceeded");
    if (S.nw == 0) w_begin = 0;
    if ((long long)h->ld + h->wcap > (long long)PARTNER_MASK)
        return fail(h, VG_EUNSUPPORTED, "too many DOs for the neighbour-list encoding");
    // tables
    S.model_off = (int)h->models_h.size();
    for (int m = 0; m < d->n_models; ++m) {
        ModelDev md;
        md.group = d->models[m].group;
        md.scene = slot;
        for (int c = 0; c < 3; ++c) md.fext[c] = d->models[m].fext[c];
        h->models_h.push_back(md);
        double2 q = make_double2(0.0, 0.0);
        if (d->models[m].shape == VG_SPHERE) {
            const double r = d->models[m].attr[0];
            const double mass = d->models[m].density * ((4.0 / 3.0 * PI) * r * r * r);
            q = make_double2(mass, 0.4 * mass * r * r);
        }
        h->mm_h.push_back(q);
    }
    S.law_off = (int)h->laws_h.size();
    for (int k = 0; k < d->n_groups * d->n_groups; ++k) {
        LawDev L;
        L.law = d->iact[k].law;
        L.pad = 0;
        for (int p = 0; p < 8; ++p) L.p[p] = d->iact[k].p[p];
        if (L.law == VG_H) {
            // ISwH's first-contact stiffness kn = K |vn|^0.4 (ISwH.cpp:90-103): the law's
            // factor K, evaluated here in the reference's left-to-right order with the
            // host's pow, so that the device evaluates one pow per first contact
            const double knh = L.p[0], meff = L.p[1], betan = L.p[4];
            L.p[6] = betan < 0.99 ? 0.957832746294131 / (1.0 - betan * betan) * pow(meff, 0.2) * pow(knh, 0.8)
                                  : 479.155951122633 * pow(meff, 0.2) * pow(knh, 0.8);
        }
        h->laws_h.push_back(L);
    }
    for (int &m : S.mdl0) m += S.model_off;
    S.w_begin = w_begin;
    for (int w = 0; w < S.nw; ++w) {
        S.walls0[w].scene = slot;
        h->wall_used[w_begin + w] = 1;
    }
    for (int &w : S.do_to_wall)
        if (w >= 0) w += w_begin;
    // tables: models (with DOModel::Analysis mass / inertia, DOModel.cpp:22-29) and laws
    const int models_cap = h->models_cap;
    rc = grow_table(h, &h->models, &h->models_cap, h->models_h);
    if (!rc && h->models_cap != models_cap) {
        if (h->mm) HIPCHK(hipFree(h->mm));
        h->mm = nullptr;
        rc = dalloc(h, &h->mm, (size_t)h->models_cap);
    }
    if (!rc) HIPCHK(hipMemcpy(h->mm, h->mm_h.data(), sizeof(double2) * h->mm_h.size(), hipMemcpyHostToDevice));
    if (!rc) rc = grow_table(h, &h->laws, &h->laws_cap, h->laws_h);
    if (rc) return rc;
    S.steps_target = steps;
    S.state = VG_SCENE_QUEUED;
    h->sc[slot] = std::move(S);
    *slot_out = slot;
    return 0;
}

// Comparator order of the 14 cell groups of a sphere's Leap traversal: its
// own cell (13) and the cells at -NB13[k] (k < 13), whose half shells hold
// it.  The comparator compares cell coordinates lexicographically in the
// axis priority `perm`, so for cells at fixed offsets from one cell it
// compares the offsets; the order, packed 4 bits per group, is the same for
// every cell of the grid.
static unsigned long long group_order(const int perm[3])
{
    static const int nb[13][3] = VG_NB13_INIT;
    int off[14][3], idx[14];
    for (int k = 0; k < 13; ++k)
        for (int c = 0; c < 3; ++c) off[k][c] = -nb[k][c];
    off[13][0] = off[13][1] = off[13][2] = 0;
    for (int k = 0; k < 14; ++k) idx[k] = k;
    std::sort(idx, idx + 14, [&](int u, int v) {
        for (int c = 0; c < 3; ++c)
            if (off[u][perm[c]] != off[v][perm[c]]) return off[u][perm[c]] < off[v][perm[c]];
        return false;
    });
    unsigned long long order = 0;
    for (int q = 0; q < 14; ++q) order |= (unsigned long long)idx[q] << (4 * q);
    return order;
}

// ---------------------------------------------------------------------------
// Rebuild boundary: place queued scenes, drop finished ones, rebuild
// ---------------------------------------------------------------------------

static int fetch_scene_state(vg_scene *h, int slot, std::vector<double> &hb);

static int place_queued(vg_scene *h)
{
    Geo &G = h->geo[h->cur];
    bool any = false;
    for (int k = 0; k < h->scap; ++k) {
        SceneHost &S = h->sc[k];
        if (S.state != VG_SCENE_QUEUED) continue;
        const int b = h->n, n = S.n, ld = h->ld;
        // scenes leaving at this boundary still hold their slots until the
        // sort: a scene that does not fit now is placed at the next boundary
        if (b + n > ld) continue;
        std::vector<unsigned> oid(n);
        for (int s = 0; s < n; ++s) oid[s] = (unsigned)s;
        HIPCHK(hipMemcpy(G.pos + b, S.pos0.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(G.vel + b, S.vel0.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(G.avl + b, S.avl0.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
        for (int c = 0; c < 6; ++c)
            HIPCHK(hipMemcpy(G.ori + (size_t)c * ld + b, S.ori0.data() + (size_t)c * n, sizeof(double) * n,
                             hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(G.oid + b, oid.data(), sizeof(unsigned) * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(G.mdl + b, S.mdl0.data(), sizeof(int) * n, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_hdr_zero, dim3(nblk(n, 256)), dim3(256), 0, h->stream, h->ell[h->ecur].hdr, b, b + n);
        if (S.nw) {
            if (!S.kin0.empty()) {
                for (int w = 0; w < S.nw; ++w)
                    if (S.kin0_set[w]) S.walls0[w].kin = 2;
                HIPCHK(hipMemcpy(h->wkin + S.w_begin, S.kin0.data(), sizeof(WallKin) * S.nw, hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(h->walls_ref + S.w_begin, S.walls0.data(), sizeof(WallDev) * S.nw,
                                 hipMemcpyHostToDevice));
            }
            HIPCHK(hipMemcpy(h->walls + S.w_begin, S.walls0.data(), sizeof(WallDev) * S.nw, hipMemcpyHostToDevice));
            for (int w = 0; w < S.nw; ++w) h->walls_h[S.w_begin + w] = S.walls0[w];
            // the hole table exists once a scene with a holed plate is placed;
            // from then on every placed scene writes its slots (zeros: no hole)
            if (S.has_hole && !h->holes) {
                if (dalloc(h, &h->holes, (size_t)h->wcap)) return VG_ENOMEM;
                HIPCHK(hipMemset(h->holes, 0, sizeof(double4) * (size_t)std::max(h->wcap, 1)));
            }
            if (h->holes)
                HIPCHK(hipMemcpy(h->holes + S.w_begin, S.holes0.data(), sizeof(double4) * S.nw,
                                 hipMemcpyHostToDevice));
            h->nw_used = std::max(h->nw_used, S.w_begin + S.nw);
        }
        if (S.nbs) h->nbs = true;
        if (S.has_pbc && !h->pbc) {
            // the traversal axes are set at every rebuild
            if (dalloc(h, &h->pbc, 1)) return VG_ENOMEM;
            h->pbc_h = S.pbc0;
            HIPCHK(hipMemcpy(h->pbc, &h->pbc_h, sizeof(PbcDev), hipMemcpyHostToDevice));
        }
        const double t2[2] = {S.t, S.t};
        HIPCHK(hipMemcpy(h->stime + 2 * k, t2, sizeof(t2), hipMemcpyHostToDevice));
        const int z3[3] = {0, 0, 0};
        HIPCHK(hipMemcpy(h->rflags + 3 * k, z3, sizeof(z3), hipMemcpyHostToDevice));
        SceneDev &D = h->sdev[k];
        D = SceneDev();
        D.dt = S.dt; D.gx = S.g[0]; D.gy = S.g[1]; D.gz = S.g[2];
        D.skin = S.skin; D.relist_half = 0.5 * S.skin * (1.0 - 1e-9);
        D.law_off = S.law_off; D.n_groups = S.n_groups;
        D.w_begin = S.w_begin; D.w_end = S.w_begin + S.nw;
        D.zon = S.zon;
        for (int c = 0; c < 3; ++c) { D.zlo[c] = S.zlo[c]; D.zhi[c] = S.zhi[c]; }
        D.active = 1;
        S.seg_begin = b;
        S.seg_end = b + n;
        S.state = VG_SCENE_RUNNING;
        S.pos0.clear(); S.vel0.clear(); S.avl0.clear(); S.ori0.clear(); S.mdl0.clear();
        S.pos0.shrink_to_fit(); S.vel0.shrink_to_fit(); S.avl0.shrink_to_fit(); S.ori0.shrink_to_fit();
        S.mdl0.shrink_to_fit();
        h->n += n;
        any = true;
    }
    if (any) update_batch_flags(h);
    return 0;
}

// LeapConsultant::Reset + subReset (LeapConsultant.cpp:108-195,
// LeapConsultant.h:57-514) for every placed scene, followed by
// SimMediator::Initiate's history restore.  Finished scenes leave the device
// here (their final state is kept on the host).
// Constrained walls: the pairs of each constrained wall in IactPairTab order
// (see k_cw_emit), the pair slot of each (sphere, wall), and the pair range
// of each wall slot, for the lists the rebuild has just made.
static int build_cw_pairs(vg_scene *h, const Geo &G, int n)
{
    hipStream_t st = h->stream;
    const int ld = h->ld;
    if (h->wcap > (1 << 13)) return fail(h, VG_EUNSUPPORTED, "constrained walls need at most 8192 wall slots");
    if (!h->cw_w) {
        if (dalloc(h, &h->cw_w, (size_t)CW_SLOTS * ld) || dalloc(h, &h->cw_p, (size_t)CW_SLOTS * ld) ||
            dalloc(h, &h->cw_cnt, (size_t)2 * (ld + 1)))
            return VG_ENOMEM;
    }
    unsigned *cnt = h->cw_cnt, *off = h->cw_cnt + (ld + 1);
    hipLaunchKernelGGL(k_cw_count, dim3(nblk(n + 1, 256)), dim3(256), 0, st, n, ld, h->ko,
                       (const unsigned *)h->ocount, (const unsigned *)h->onbr, (const WallDev *)h->walls, cnt,
                       h->flags);
    size_t need = 0;
    (void)rocprim::exclusive_scan(nullptr, need, cnt, off, 0u, (size_t)n + 1, rocprim::plus<unsigned>(), st);
    if (ensure_tmp(h, need)) return VG_EHIP;
    HIPCHK(rocprim::exclusive_scan(h->tmp, need, cnt, off, 0u, (size_t)n + 1, rocprim::plus<unsigned>(), st));
    unsigned np = 0;
    HIPCHK(hipMemcpyAsync(&np, off + n, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if ((long long)np > h->cw_cap) {
        auto F = [](void *p) {
            if (p) (void)hipFree(p);
        };
        F(h->cw_val); F(h->cw_stamp); F(h->cw_key[0]); F(h->cw_key[1]); F(h->cw_vidx[0]); F(h->cw_vidx[1]);
        h->cw_val = nullptr; h->cw_stamp = nullptr;
        h->cw_key[0] = h->cw_key[1] = nullptr; h->cw_vidx[0] = h->cw_vidx[1] = nullptr;
        h->cw_cap = (long long)np + np / 2 + 1024;
        const size_t c = (size_t)h->cw_cap;
        if (dalloc(h, &h->cw_val, 3 * c) || dalloc(h, &h->cw_stamp, c) || dalloc(h, &h->cw_key[0], c) ||
            dalloc(h, &h->cw_key[1], c) || dalloc(h, &h->cw_vidx[0], c) || dalloc(h, &h->cw_vidx[1], c))
            return VG_ENOMEM;
    }
    hipLaunchKernelGGL(k_fill_int, dim3(nblk((long long)CW_SLOTS * ld, 256)), dim3(256), 0, st, h->cw_w,
                       (long long)CW_SLOTS * ld, -1);
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(h->wcap, 256)), dim3(256), 0, st, h->cw_beg, (long long)h->wcap, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(h->wcap, 256)), dim3(256), 0, st, h->cw_end, (long long)h->wcap, 0);
    if (np > 0) {
        hipLaunchKernelGGL(k_cw_emit, dim3(nblk(n, 256)), dim3(256), 0, st, n, ld, h->ko, (const unsigned *)h->ocount,
                           (const unsigned *)h->onbr, (const WallDev *)h->walls, G, (const ModelDev *)h->models,
                           (const SceneDev *)h->scenes, (const unsigned *)off, h->cw_key[0], h->cw_vidx[0]);
        need = 0;
        (void)rocprim::radix_sort_pairs(nullptr, need, h->cw_key[0], h->cw_key[1], h->cw_vidx[0], h->cw_vidx[1],
                                        (size_t)np, 0, 64, st);
        if (ensure_tmp(h, need)) return VG_EHIP;
        HIPCHK(rocprim::radix_sort_pairs(h->tmp, need, h->cw_key[0], h->cw_key[1], h->cw_vidx[0], h->cw_vidx[1],
                                         (size_t)np, 0, 64, st));
        hipLaunchKernelGGL(k_cw_finalize, dim3(nblk(np, 256)), dim3(256), 0, st, (long long)np,
                           (const unsigned long long *)h->cw_key[1], (const unsigned *)h->cw_vidx[1], ld, h->cw_w,
                           h->cw_p, h->cw_stamp, h->cw_beg, h->cw_end);
    }
    h->cw_n = np;
    HIPCHK(hipGetLastError());
    return 0;
}

static int export_active(vg_scene *h, int slot, ActiveOut *out, size_t cap, unsigned long long *dcnt,
                         unsigned *d_sid, unsigned *d_wid, bool backup);
static int id_maps(vg_scene *h, int slot, unsigned **d_sid, unsigned **d_wid);

// A scene that loses spheres to its zone of interest at a rebuild: its
// records re-keyed by ModifyPair, restored over the rebuilt lists.
struct ZoiPlan {
    int slot = -1;
    unsigned long long *keys = nullptr;   // sorted new keys (of equal keys the last applies)
    RestoreRec *recs = nullptr;
    long long nrec = 0;
};

static void zoi_free(std::vector<ZoiPlan> &plans)
{
    for (ZoiPlan &p : plans) {
        if (p.keys) (void)hipFree(p.keys);
        if (p.recs) (void)hipFree(p.recs);
        p.keys = nullptr;
        p.recs = nullptr;
    }
}

// Consultant::CleanUp (Consultant.cpp:264-319) at a rebuild boundary, after
// RebuildIactRecordTab and before Reset (LeapConsultant.cpp:58-68), and at a
// scene's placement (SimMediator.cpp:122-134): spheres of stepping scenes
// outside their zone of interest are erased.  Their scene's records (the
// active pairs of the lists about to be replaced) move with ModifyPair; the
// DO ids and ranks close up; the erased spheres take ranks in the dropped
// region of the rebuild's sort.  The host maps of DO ids follow.
static int zoi_prepare(vg_scene *h, std::vector<ZoiPlan> &plans)
{
    hipStream_t st = h->stream;
    const int n = h->n, ld = h->ld;
    if (upload_scenes(h)) return VG_EHIP;   // scenes placed at this boundary carry their zone
    if (!h->zoi_dead) {
        if (dalloc(h, &h->zoi_dead, (size_t)ld) || dalloc(h, &h->zoi_cnt, (size_t)h->scap) ||
            dalloc(h, &h->zoi_rank, (size_t)ld + 1) || dalloc(h, &h->zoi_pre, (size_t)ld + 1))
            return VG_ENOMEM;
    }
    HIPCHK(hipMemsetAsync(h->zoi_cnt, 0, sizeof(int) * h->scap, st));
    hipLaunchKernelGGL(k_zoi_mark, dim3(nblk(n, 256)), dim3(256), 0, st, h->geo[h->cur], n,
                       (const SceneDev *)h->scenes, (const ModelDev *)h->models, h->zoi_dead, h->zoi_cnt);
    std::vector<int> cnt(h->scap);
    HIPCHK(hipMemcpyAsync(cnt.data(), h->zoi_cnt, sizeof(int) * h->scap, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int> dead_slots;
    for (int k = 0; k < h->scap; ++k)
        if (cnt[k] > 0 && h->sc[k].state == VG_SCENE_RUNNING) dead_slots.push_back(k);
    if (dead_slots.empty()) return 0;
    size_t need = 0;
    for (int k : dead_slots) {
        SceneHost &S = h->sc[k];
        const int b = S.seg_begin, e = S.seg_end, n_old = S.n;
        ZoiPlan plan;
        plan.slot = k;
        // the scene's records, keyed by its DO ids before the erase (no lists yet at its placement)
        ActiveOut *recs = nullptr;
        long long m = 0;
        if (S.steps_done > 0) {
            unsigned long long *dcnt = nullptr;
            unsigned *d_sid = nullptr, *d_wid = nullptr;
            if (dalloc(h, &dcnt, 1) || id_maps(h, k, &d_sid, &d_wid)) return VG_ENOMEM;
            if (export_active(h, k, nullptr, 0, dcnt, d_sid, d_wid, false)) return VG_EHIP;
            unsigned long long c = 0;
            HIPCHK(hipMemcpyAsync(&c, dcnt, sizeof(c), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            m = (long long)c;
            if (m > 0) {
                if (dalloc(h, &recs, (size_t)m)) return VG_ENOMEM;
                if (export_active(h, k, recs, (size_t)m, dcnt, d_sid, d_wid, false)) return VG_EHIP;
            }
            HIPCHK(hipStreamSynchronize(st));
            (void)hipFree(dcnt); (void)hipFree(d_sid); (void)hipFree(d_wid);
        }
        // erase flags by rank and their prefix count
        hipLaunchKernelGGL(k_zoi_rank, dim3(nblk(e - b, 256)), dim3(256), 0, st, h->geo[h->cur], b, e,
                           (const unsigned char *)h->zoi_dead, h->zoi_rank);
        HIPCHK(hipMemsetAsync(h->zoi_rank + e, 0, sizeof(unsigned), st));
        need = 0;
        (void)rocprim::exclusive_scan(nullptr, need, h->zoi_rank + b, h->zoi_pre + b, 0u, (size_t)n_old + 1,
                                      rocprim::plus<unsigned>(), st);
        if (ensure_tmp(h, need)) return VG_EHIP;
        HIPCHK(rocprim::exclusive_scan(h->tmp, need, h->zoi_rank + b, h->zoi_pre + b, 0u, (size_t)n_old + 1,
                                       rocprim::plus<unsigned>(), st));
        std::vector<unsigned> dead_by_rank((size_t)n_old);
        HIPCHK(hipMemcpyAsync(dead_by_rank.data(), h->zoi_rank + b, sizeof(unsigned) * n_old, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        // DO ids close up (DOWorld::EraseDOStatus); ModifyPair's map of every DO id:
        // id -> id + jump, jump = minus the erased DOs up to and including id
        std::vector<char> dead_do((size_t)S.n_do, 0);
        for (int r = 0; r < n_old; ++r)
            if (dead_by_rank[(size_t)r]) dead_do[S.sphere_ids[(size_t)r]] = 1;
        std::vector<unsigned> jump((size_t)S.n_do);
        unsigned ul_dead = 0;
        for (int id = 0; id < S.n_do; ++id) {
            if (dead_do[(size_t)id]) --ul_dead;
            jump[(size_t)id] = (unsigned)id + ul_dead;
        }
        if (m > 0) {
            unsigned *d_jump = nullptr, *prio_src = nullptr, *prio = nullptr, *idx[2] = {nullptr, nullptr};
            unsigned long long *okey = nullptr, *nkey = nullptr, *k2[2] = {nullptr, nullptr};
            RestoreRec *rr = nullptr;
            const size_t M = (size_t)m;
            if (dalloc(h, &d_jump, (size_t)S.n_do) || dalloc(h, &prio_src, M) || dalloc(h, &prio, M) ||
                dalloc(h, &idx[0], M) || dalloc(h, &idx[1], M) || dalloc(h, &okey, M) || dalloc(h, &nkey, M) ||
                dalloc(h, &k2[0], M) || dalloc(h, &k2[1], M) || dalloc(h, &rr, M) || dalloc(h, &plan.keys, M) ||
                dalloc(h, &plan.recs, M))
                return VG_ENOMEM;
            HIPCHK(hipMemcpyAsync(d_jump, jump.data(), sizeof(unsigned) * S.n_do, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_zoi_records, dim3(nblk(m, 256)), dim3(256), 0, st, (const ActiveOut *)recs, m,
                               (const unsigned *)d_jump, okey, nkey, prio_src, rr);
            // the table's (std::map) order: records by old key
            need = 0;
            (void)rocprim::radix_sort_pairs(nullptr, need, okey, k2[0], prio_src, idx[0], M, 0, 64, st);
            if (ensure_tmp(h, need)) return VG_EHIP;
            HIPCHK(rocprim::radix_sort_pairs(h->tmp, need, okey, k2[0], prio_src, idx[0], M, 0, 64, st));
            hipLaunchKernelGGL(k_zoi_prio, dim3(nblk(m, 256)), dim3(256), 0, st, (const unsigned long long *)k2[0],
                               (const unsigned *)idx[0], (const unsigned long long *)nkey, m, prio);
            // order by (new key, priority): by priority, then stably by new key
            hipLaunchKernelGGL(k_iota, dim3(nblk(m, 256)), dim3(256), 0, st, prio_src, m);
            need = 0;
            (void)rocprim::radix_sort_pairs(nullptr, need, prio, (unsigned *)k2[1], prio_src, idx[1], M, 0, 32, st);
            if (ensure_tmp(h, need)) return VG_EHIP;
            HIPCHK(rocprim::radix_sort_pairs(h->tmp, need, prio, (unsigned *)k2[1], prio_src, idx[1], M, 0, 32, st));
            hipLaunchKernelGGL(k_gather_u64, dim3(nblk(m, 256)), dim3(256), 0, st, (const unsigned long long *)nkey,
                               (const unsigned *)idx[1], m, k2[0]);
            need = 0;
            (void)rocprim::radix_sort_pairs(nullptr, need, k2[0], plan.keys, idx[1], idx[0], M, 0, 64, st);
            if (ensure_tmp(h, need)) return VG_EHIP;
            HIPCHK(rocprim::radix_sort_pairs(h->tmp, need, k2[0], plan.keys, idx[1], idx[0], M, 0, 64, st));
            hipLaunchKernelGGL(k_gather_restore, dim3(nblk(m, 256)), dim3(256), 0, st, (const RestoreRec *)rr,
                               (const unsigned *)idx[0], m, plan.recs);
            plan.nrec = m;
            HIPCHK(hipStreamSynchronize(st));
            for (void *q : {(void *)d_jump, (void *)prio_src, (void *)prio, (void *)idx[0], (void *)idx[1],
                            (void *)okey, (void *)nkey, (void *)k2[0], (void *)k2[1], (void *)rr, (void *)recs})
                if (q) (void)hipFree(q);
        } else if (recs) {
            (void)hipFree(recs);
        }
        // the host maps of the closed-up DO ids
        std::vector<unsigned> sid_new;
        sid_new.reserve((size_t)(n_old - cnt[k]));
        for (int r = 0; r < n_old; ++r)
            if (!dead_by_rank[(size_t)r]) sid_new.push_back(jump[S.sphere_ids[(size_t)r]]);
        for (unsigned &w : S.wall_ids) w = jump[w];
        std::vector<int> model_new;
        model_new.reserve((size_t)(S.n_do - cnt[k]));
        for (int id = 0; id < S.n_do; ++id)
            if (!dead_do[(size_t)id]) model_new.push_back(S.do_model[(size_t)id]);
        S.do_model.swap(model_new);
        S.n_do -= cnt[k];
        S.n = (int)sid_new.size();
        S.sphere_ids.swap(sid_new);
        S.do_to_sphere.assign((size_t)S.n_do, -1);
        S.do_to_wall.assign((size_t)S.n_do, -1);
        for (int r = 0; r < S.n; ++r) S.do_to_sphere[S.sphere_ids[(size_t)r]] = r;
        for (int w = 0; w < S.nw; ++w) S.do_to_wall[S.wall_ids[(size_t)w]] = S.w_begin + w;
        plans.push_back(plan);
    }
    // ranks close up; erased spheres go to the dropped region after the live and retiring scenes
    long long dead_base = 0;
    for (int k = 0; k < h->scap; ++k) {
        const SceneHost &S = h->sc[k];
        if (S.state == VG_SCENE_RUNNING || S.state == VG_SCENE_FINISHED || S.state == VG_SCENE_RETIRING)
            dead_base += S.n;
    }
    for (int k : dead_slots) {
        const SceneHost &S = h->sc[k];
        hipLaunchKernelGGL(k_zoi_newoid, dim3(nblk(S.seg_end - S.seg_begin, 256)), dim3(256), 0, st,
                           h->geo[h->cur], S.seg_begin, S.seg_end, (const unsigned char *)h->zoi_dead,
                           (const unsigned *)h->zoi_pre, (unsigned)dead_base);
        dead_base += cnt[k];
    }
    HIPCHK(hipGetLastError());
    return 0;
}

// Candidate list of a periodic scene (one scene per handle): the occupied
// cells sorted into comparator order, their pairs counted and laid down in
// the reference's order (k_pbc_count / k_pbc_emit), then spread to both
// spheres' lists by a stable sort on the sphere (k_pbc_entries ..
// k_pbc_lists).  Fills ocount / onbr and flags[5] as k_outer does.
static int pbc_candidates(vg_scene *h, const Geo &G, int n, const std::vector<GridDev> &grids)
{
    hipStream_t st = h->stream;
    if ((size_t)n + 1 > h->pbc_cells_cap) {
        for (auto p : {(void *)h->pbc_ckey[0], (void *)h->pbc_ckey[1], (void *)h->pbc_cfirst[0],
                       (void *)h->pbc_cfirst[1], (void *)h->pbc_pcount, (void *)h->pbc_poff})
            if (p) HIPCHK(hipFree(p));
        h->pbc_cells_cap = (size_t)n + 1 + (size_t)n / 4;
        const size_t c = h->pbc_cells_cap;
        if (dalloc(h, &h->pbc_ckey[0], c) || dalloc(h, &h->pbc_ckey[1], c) || dalloc(h, &h->pbc_cfirst[0], c) ||
            dalloc(h, &h->pbc_cfirst[1], c) || dalloc(h, &h->pbc_pcount, c) || dalloc(h, &h->pbc_poff, c))
            return VG_ENOMEM;
    }
    if (!h->pbc_ncell && dalloc(h, &h->pbc_ncell, 1)) return VG_ENOMEM;
    if (!h->pbc_first && dalloc(h, &h->pbc_first, (size_t)h->ld)) return VG_ENOMEM;
    HIPCHK(hipMemsetAsync(h->pbc_ncell, 0, sizeof(unsigned), st));
    hipLaunchKernelGGL(k_pbc_cells, dim3(nblk(n, 256)), dim3(256), 0, st, G, n, (const int *)h->lin,
                       (const SceneDev *)h->scenes, h->pbc_ckey[0], h->pbc_cfirst[0], h->pbc_ncell);
    unsigned nc = 0;
    HIPCHK(hipMemcpyAsync(&nc, h->pbc_ncell, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    size_t need = 0;
    (void)rocprim::radix_sort_pairs(nullptr, need, h->pbc_ckey[0], h->pbc_ckey[1], h->pbc_cfirst[0],
                                    h->pbc_cfirst[1], (size_t)nc, 0, 63, st);
    if (ensure_tmp(h, need)) return VG_EHIP;
    HIPCHK(rocprim::radix_sort_pairs(h->tmp, need, h->pbc_ckey[0], h->pbc_ckey[1], h->pbc_cfirst[0],
                                     h->pbc_cfirst[1], (size_t)nc, 0, 63, st));
    PbcArgs pa;
    pa.g = G; pa.n = n; pa.ld = h->ld; pa.scene = h->scenes; pa.cstart = h->cstart; pa.cend = h->cend;
    pa.safe = h->safe; pa.walls = h->walls; pa.holes = h->holes; pa.pbc = h->pbc; pa.t = tables(h);
    pa.ckey = h->pbc_ckey[1]; pa.cfirst = h->pbc_cfirst[1]; pa.ncell = (int)nc;
    pa.pcount = h->pbc_pcount; pa.poff = h->pbc_poff; pa.pairs = nullptr;
    pa.nbs = h->nbs ? 1 : 0;
    pa.nc[0] = (unsigned)(grids[0].nx - 3); pa.nc[1] = (unsigned)(grids[0].ny - 3); pa.nc[2] = (unsigned)(grids[0].nz - 3);
    HIPCHK(hipMemsetAsync(h->pbc_pcount + nc, 0, sizeof(unsigned), st));
    if (nc) hipLaunchKernelGGL(k_pbc_count, dim3(nblk(nc, 64)), dim3(64), 0, st, pa);
    need = 0;
    (void)rocprim::exclusive_scan(nullptr, need, h->pbc_pcount, h->pbc_poff, 0u, (size_t)nc + 1,
                                  rocprim::plus<unsigned>(), st);
    if (ensure_tmp(h, need)) return VG_EHIP;
    HIPCHK(rocprim::exclusive_scan(h->tmp, need, h->pbc_pcount, h->pbc_poff, 0u, (size_t)nc + 1,
                                   rocprim::plus<unsigned>(), st));
    unsigned np = 0;
    HIPCHK(hipMemcpyAsync(&np, h->pbc_poff + nc, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (np >= (1u << 30)) return fail(h, VG_ECAPACITY, "periodic candidate list: too many pairs");
    if (np > h->pbc_pairs_cap) {
        for (auto p : {(void *)h->pbc_pairs, (void *)h->pbc_ekey[0], (void *)h->pbc_ekey[1], (void *)h->pbc_eval[0],
                       (void *)h->pbc_eval[1]})
            if (p) HIPCHK(hipFree(p));
        h->pbc_pairs_cap = (size_t)np + (size_t)np / 4 + 64;
        const size_t c = h->pbc_pairs_cap;
        if (dalloc(h, &h->pbc_pairs, c) || dalloc(h, &h->pbc_ekey[0], 2 * c) || dalloc(h, &h->pbc_ekey[1], 2 * c) ||
            dalloc(h, &h->pbc_eval[0], 2 * c) || dalloc(h, &h->pbc_eval[1], 2 * c))
            return VG_ENOMEM;
    }
    pa.pairs = h->pbc_pairs;
    if (nc) hipLaunchKernelGGL(k_pbc_emit, dim3(nblk(nc, 64)), dim3(64), 0, st, pa);
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(n, 256)), dim3(256), 0, st, reinterpret_cast<int *>(h->ocount),
                       (long long)n, 0);
    const unsigned ne = 2 * np;
    if (np) {
        hipLaunchKernelGGL(k_pbc_entries, dim3(nblk(np, 256)), dim3(256), 0, st, (const uint2 *)h->pbc_pairs, np,
                           h->ld, h->pbc_ekey[0], h->pbc_eval[0]);
        need = 0;
        (void)rocprim::radix_sort_pairs(nullptr, need, h->pbc_ekey[0], h->pbc_ekey[1], h->pbc_eval[0],
                                        h->pbc_eval[1], (size_t)ne, 0, 32, st);
        if (ensure_tmp(h, need)) return VG_EHIP;
        HIPCHK(rocprim::radix_sort_pairs(h->tmp, need, h->pbc_ekey[0], h->pbc_ekey[1], h->pbc_eval[0],
                                         h->pbc_eval[1], (size_t)ne, 0, 32, st));
        hipLaunchKernelGGL(k_pbc_segstart, dim3(nblk(ne, 256)), dim3(256), 0, st, (const unsigned *)h->pbc_ekey[1],
                           ne, h->pbc_first);
        hipLaunchKernelGGL(k_pbc_lists, dim3(nblk(ne, 256)), dim3(256), 0, st, (const unsigned *)h->pbc_ekey[1],
                           (const unsigned *)h->pbc_eval[1], ne, (const uint2 *)h->pbc_pairs,
                           (const unsigned *)h->pbc_first, h->ld, h->ko, h->ki, h->ocount, h->onbr, h->flags);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

static void free_ell(Ell &e);
static int alloc_ell(vg_scene *h, Ell &e, int n, int ki);

static int rebuild(vg_scene *h)
{
    hipStream_t st = h->stream;
    PhaseClock pc;
    if (pc.on) {
        // debugging: the steps queued since the last rebuild, timed apart from the rebuild
        HIPCHK(hipStreamSynchronize(st));
        pc.mark("steps before (synced)");
    }
    // finished scenes: keep the result, retire
    bool retiring = false;
    for (int k = 0; k < h->scap; ++k) {
        SceneHost &S = h->sc[k];
        if (S.state == VG_SCENE_FINISHED && h->batch) {
            if (fetch_scene_state(h, k, S.result)) return VG_EHIP;
            S.state = VG_SCENE_RETIRING;
        }
        if (S.state == VG_SCENE_RETIRING) {
            h->sdev[k].active = -1;
            retiring = true;
        }
    }
    pc.mark("finished / retiring");
    if (place_queued(h)) return VG_EHIP;
    pc.mark("place queued");
    std::vector<ZoiPlan> zplans;
    struct ZoiGuard {
        std::vector<ZoiPlan> &p;
        ~ZoiGuard() { zoi_free(p); }
    } zguard{zplans};
    if (h->any_zoi && h->n > 0) {
        const int rc = zoi_prepare(h, zplans);
        if (rc) return rc;
        if (!zplans.empty()) {
            // the pre-rebuild snapshot and per-entry records name erased DOs and old ids
            h->snap_valid = false;
            h->ent_valid = false;
        }
    }
    const int n_total = h->n;
    // rank offsets: live scenes by slot, then the retiring ones
    int n_live = 0, max_slot = -1;
    std::vector<int> live;
    for (int k = 0; k < h->scap; ++k) {
        SceneHost &S = h->sc[k];
        if (S.state == VG_SCENE_RUNNING || S.state == VG_SCENE_FINISHED) {
            h->sdev[k].rank_off = n_live;
            n_live += S.n;
            live.push_back(k);
            max_slot = k;
        }
    }
    {
        int r = n_live;
        for (int k = 0; k < h->scap; ++k)
            if (h->sc[k].state == VG_SCENE_RETIRING) {
                h->sdev[k].rank_off = r;
                r += h->sc[k].n;
            }
    }
    if (upload_scenes(h)) return VG_EHIP;
    pc.mark("rank offsets + upload");
    Geo &G0 = h->geo[h->cur];
    Geo &G1 = h->geo[h->cur ^ 1];
    const int S_live = (int)live.size();
    if (n_total > 0) {
        // bounds and quantile sorts over the current layout
        HIPCHK(hipMemsetAsync(h->red, 0, sizeof(unsigned long long) * 8 * h->scap, st));
        hipLaunchKernelGGL(k_bounds, dim3(nblk(n_total, 256)), dim3(256), 0, st, G0, n_total,
                           (const SceneDev *)h->scenes, (const ModelDev *)h->models, h->red, h->inv, h->xs[0],
                           h->ys[0], h->zs[0], h->sid[0], (unsigned)(max_slot + 1));
    }
    std::vector<GridDev> grids(h->scap);
    std::vector<long long> cell_off(h->scap, 0);
    long long cells = 0;
    if (S_live > 0) {
        // 2%-trimmed spans pick the comparator (LeapConsultant.cpp:139-191):
        // the elements the reference reads from each sorted coordinate,
        // found by radix selection over every live scene at once
        std::vector<QselTarget> tg((size_t)6 * S_live);
        std::vector<int> qmap(h->scap, -1);
        const double discard = 0.02, keep = 1.0 - discard;
        for (int q = 0; q < S_live; ++q) {
            const SceneHost &S = h->sc[live[q]];
            const double n = (double)S.n;
            qmap[live[q]] = q;
            // (size_t)(discard*n) and (size_t)(keep*n - 1) as the reference indexes, clamped to the scene
            const long long ilo = std::max(0ll, std::min((long long)(discard * n), (long long)S.n - 1));
            const long long ihi = std::max(0ll, std::min((long long)std::floor(keep * n - 1.0), (long long)S.n - 1));
            for (int c = 0; c < 3; ++c) {
                tg[(size_t)6 * q + 2 * c] = QselTarget{0ull, (unsigned long long)ilo};
                tg[(size_t)6 * q + 2 * c + 1] = QselTarget{0ull, (unsigned long long)ihi};
            }
        }
        HIPCHK(hipMemcpyAsync(h->qsel_tg, tg.data(), sizeof(QselTarget) * tg.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(h->qmap, qmap.data(), sizeof(int) * qmap.size(), hipMemcpyHostToDevice, st));
        const long long nh = (long long)tg.size() * 256;
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(nh, 256)), dim3(256), 0, st, reinterpret_cast<int *>(h->qsel_hist), nh,
                           0);
        for (int pass = 0; pass < 8; ++pass) {
            hipLaunchKernelGGL(k_qsel_hist, dim3(nblk(n_total, 256)), dim3(256), 0, st, G0, n_total,
                               (const unsigned *)h->sid[0], (const int *)h->qmap, h->scap,
                               (const QselTarget *)h->qsel_tg, pass, h->qsel_hist);
            hipLaunchKernelGGL(k_qsel_pick, dim3(nblk((long long)tg.size(), 4)), dim3(256), 0, st, h->qsel_tg,
                               (int)tg.size(), pass, h->qsel_hist, h->qval);
        }
        pc.mark("quantile selection launch");
        std::vector<unsigned long long> red((size_t)8 * h->scap);
        std::vector<double> q6((size_t)6 * S_live);
        HIPCHK(hipMemcpyAsync(red.data(), h->red, sizeof(unsigned long long) * red.size(), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(q6.data(), h->qval, sizeof(double) * q6.size(), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        pc.mark("sorts + readback");
        for (int q = 0; q < S_live; ++q) {
            const int k = live[q];
            const SceneHost &S = h->sc[k];
            const unsigned long long *rr = &red[(size_t)8 * k];
            double xmax = from_ord(rr[0]), ymax = from_ord(rr[1]), zmax = from_ord(rr[2]);
            const double xmin = from_ord(~rr[3]), ymin = from_ord(~rr[4]), zmin = from_ord(~rr[5]);
            const double vmax = from_ord(rr[6]);
            double rmax = from_ord(rr[7]);
            if (S.nbs) {
                // NBSConsultant::Reset: maxima start from numeric_limits<double>::min() and the
                // radius is the last local DO's range (NBSConsultant.cpp:106-141)
                xmax = std::max(xmax, DBL_MIN); ymax = std::max(ymax, DBL_MIN); zmax = std::max(zmax, DBL_MIN);
                rmax = S.nbs_range;
            }
            const double *qq = &q6[(size_t)6 * q];
            const double xspan = qq[1] - qq[0], yspan = qq[3] - qq[2], zspan = qq[5] - qq[4];
            GridDev &gd = grids[k];
            // TrirLeX (x,y,z), TrirLeY (y,x,z), TrirLeZ (z,x,y) (NBSConsultant.h:52-118)
            int perm[3];
            if (S.nbs) {
                // NBS visits its zones by index, x outermost (NBSConsultant.cpp:197-273)
                perm[0] = 0; perm[1] = 1; perm[2] = 2;
            } else if (xspan >= yspan) {
                if (xspan >= zspan) { perm[0] = 0; perm[1] = 1; perm[2] = 2; }
                else { perm[0] = 2; perm[1] = 0; perm[2] = 1; }
            } else {
                if (yspan >= zspan) { perm[0] = 1; perm[1] = 0; perm[2] = 2; }
                else { perm[0] = 2; perm[1] = 0; perm[2] = 1; }
            }
            const double dtu = h->up_iact * S.dt;
            const double glen = olen(S.g);
            const double zr = S.sf * (2.0 * rmax + dtu * vmax + 0.5 * dtu * dtu * glen);
            if (h->skin_auto) {
                // twice the fastest sphere's path over the coming rebuild period (with a
                // margin), so that a relist between rebuilds stays the exception
                double sk = 2.2 * (dtu * vmax + 0.5 * dtu * dtu * glen);
                sk = std::min(std::max(sk, h->skin_floor), h->skin_cap);
                sk = std::max(sk, S.bond_reach * 1.01);
                h->sc[k].skin = sk;
                h->sdev[k].skin = sk;
                h->sdev[k].relist_half = 0.5 * sk * (1.0 - 1e-9);
            }
            unsigned ncx = (unsigned)std::ceil((xmax - xmin) / zr); ncx = std::max(ncx, 1u);
            unsigned ncy = (unsigned)std::ceil((ymax - ymin) / zr); ncy = std::max(ncy, 1u);
            unsigned ncz = (unsigned)std::ceil((zmax - zmin) / zr); ncz = std::max(ncz, 1u);
            if (ncx >= (1u << 17) || ncy >= (1u << 17) || ncz >= (1u << 17))
                return fail(h, VG_EUNSUPPORTED, "cell grid exceeds 2^17 cells per axis");
            gd.xmin = xmin; gd.ymin = ymin; gd.zmin = zmin; gd.zr = zr;
            gd.safe_v = dtu;
            gd.safe_g = 0.5 * dtu * dtu * glen;
            gd.sf = S.sf;
            gd.nx = (int)ncx + 3; gd.ny = (int)ncy + 3; gd.nz = (int)ncz + 3;
            gd.perm[0] = perm[0]; gd.perm[1] = perm[1]; gd.perm[2] = perm[2];
            gd.order = group_order(perm);
            if (h->pbc && S.has_pbc) {
                // the axes the Leap traversal wraps (LeapConsultant.h:146-153)
                const double span[3] = {xmax - xmin, ymax - ymin, zmax - zmin};
                const unsigned ncl[3] = {ncx, ncy, ncz};
                for (int c = 0; c < 3; ++c) {
                    h->pbc_h.pax[c] = h->pbc_h.on[c] && span[c] >= 0.5 * h->pbc_h.len[c];
                    h->pbc_h.maxid[c] = ncl[c] - 1u;
                }
                HIPCHK(hipMemcpyAsync(h->pbc, &h->pbc_h, sizeof(PbcDev), hipMemcpyHostToDevice, st));
            }
            const long long c = (long long)gd.nx * gd.ny * gd.nz;
            cell_off[k] = cells;
            cells += c;
        }
        if (h->batch) {
            // a scene whose fastest sphere crosses more than its largest radius
            // in one step has blown up (contact detection cannot follow it): it
            // leaves the batch (failed) and the rebuild starts over without it;
            // the other scenes are not affected
            bool failed = false;
            for (int q = 0; q < S_live; ++q) {
                const int k = live[q];
                const unsigned long long *rr = &red[(size_t)8 * k];
                const double vmax = from_ord(rr[6]), rmax = from_ord(rr[7]);
                if (!(vmax * h->sc[k].dt <= rmax)) {
                    h->sc[k].state = VG_SCENE_RETIRING;
                    h->sc[k].failed = true;
                    h->sdev[k].active = -1;
                    failed = true;
                    if (getenv("VG_DEBUG"))
                        fprintf(stderr, "[vg] scene %d failed: vmax %.3e m/s, rmax %.3e m\n", k, vmax, rmax);
                }
            }
            if (failed) return rebuild(h);
        }
        if (getenv("VG_DEBUG")) {
            for (int q = 0; q < S_live; ++q) {
                const int k = live[q];
                const unsigned long long *rr = &red[(size_t)8 * k];
                const GridDev &gd = grids[k];
                fprintf(stderr, "[vg] rebuild %lld scene %d n %d vmax %.3e rmax %.3e zr %.3e cells %dx%dx%d\n",
                        h->rebuilds, k, h->sc[k].n, from_ord(rr[6]), from_ord(rr[7]), gd.zr, gd.nx, gd.ny, gd.nz);
            }
        }
        if (cells > (1ll << 31)) return fail(h, VG_EUNSUPPORTED, "cell grid too large (sparse domain)");
        for (int q = 0; q < S_live; ++q) {
            const int k = live[q];
            h->sdev[k].G = grids[k];
            h->sdev[k].cell_off = cell_off[k];
        }
        if (upload_scenes(h)) return VG_EHIP;
        if (cells > h->cells_cap) {
            if (h->cstart) HIPCHK(hipFree(h->cstart));
            if (h->cend) HIPCHK(hipFree(h->cend));
            h->cstart = h->cend = nullptr;
            h->cells_cap = cells + cells / 4;
            if (dalloc(h, &h->cstart, h->cells_cap) || dalloc(h, &h->cend, h->cells_cap)) return VG_ENOMEM;
        }
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(cells, 256)), dim3(256), 0, st, h->cstart, cells, 0);
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(cells, 256)), dim3(256), 0, st, h->cend, cells, 0);
    }
    if (n_total > 0) {
        // stable sort by (scene slot, Hilbert cell key): ranks are in DO-id order
        // within each scene, so each cell's spheres stay in id order
        int slot_bits = 0;
        while ((1 << slot_bits) <= max_slot + 1) ++slot_bits;
        // cell coordinates are at most nc (the grid's cells per axis without
        // its margin); below 2^b their Hilbert index lies below 2^(3b) (the
        // curve fills the corner cube [0, 2^b)^3 first), so the scene slot sits
        // right above the largest grid's 3b bits and the sort skips the index
        // bits that are zero in every key; the order is the same
        int b = 1;
        for (int k : live) {
            const GridDev &gd = grids[k];
            const int m = std::max(gd.nx, std::max(gd.ny, gd.nz)) - 2;   // nc + 1
            while ((1 << b) < m) ++b;
        }
        const int hbits = 3 * b;
        const unsigned long long retired_key = (unsigned long long)(max_slot + 1) << hbits;
        const int end_bit = std::min(64, hbits + slot_bits);
        hipLaunchKernelGGL(k_cellkey, dim3(nblk(n_total, 256)), dim3(256), 0, st, G0, n_total,
                           (const SceneDev *)h->scenes, (const ModelDev *)h->models, (const unsigned *)h->inv,
                           retired_key, hbits, h->keys, h->svals);
        size_t need = 0;
        (void)rocprim::radix_sort_pairs(nullptr, need, h->keys, h->skeys, h->svals, h->perm, (size_t)n_total, 0,
                                        end_bit, st);
        if (ensure_tmp(h, need)) return VG_EHIP;
        HIPCHK(rocprim::radix_sort_pairs(h->tmp, need, h->keys, h->skeys, h->svals, h->perm, (size_t)n_total, 0,
                                         end_bit, st));
    }
    pc.mark("grids + cell sort launch");
    const int n = n_live;
    // the permuted rank / model arrays go to the spares, which both buffers
    // then share; the old ones stay readable through G0 for the history carry
    Geo G1w = G1;
    G1w.oid = h->oid_spare;
    G1w.mdl = h->mdl_spare;
    if (n > 0) {
        hipLaunchKernelGGL(k_reorder, dim3(nblk(n, 256)), dim3(256), 0, st, G0, G1w, n, h->ld, (const unsigned *)h->perm,
                           (const SceneDev *)h->scenes, (const ModelDev *)h->models, h->safe, h->lin,
                           h->rec_valid ? (const double *)h->rec_imp : (const double *)nullptr, h->rec_imp2);
        if (h->rec_valid) std::swap(h->rec_imp, h->rec_imp2);
        hipLaunchKernelGGL(k_cellstart, dim3(nblk(n, 256)), dim3(256), 0, st, h->lin, n, h->cstart, h->cend);
        // candidate list in one pass; grow the capacity and redo if a sphere overflowed
        OuterArgs oa;
        oa.g = G1w; oa.n = n; oa.ld = h->ld; oa.scenes = h->scenes; oa.cstart = h->cstart; oa.cend = h->cend;
        oa.safe = h->safe; oa.walls = h->walls; oa.t = tables(h); oa.count = h->ocount; oa.nbr = h->onbr;
        oa.ko = h->ko; oa.ki = h->ki; oa.single_group = h->single_group; oa.flags = h->flags;
        oa.ovf = h->batch ? h->ovf : nullptr;
        oa.holes = h->holes;
        if (h->batch) HIPCHK(hipMemsetAsync(h->ovf, 0, sizeof(int) * h->scap, st));
        int need = 0;   // the longest neighbour list the candidates make, where it exceeds the capacity
        for (;;) {
            HIPCHK(hipMemsetAsync(h->flags + 5, 0, 2 * sizeof(int), st));
            if (h->pbc || h->nbs) {
                if (int rc = pbc_candidates(h, G1w, n, grids)) return rc;
            } else if (h->holes) {
                hipLaunchKernelGGL(k_outer<true>, dim3(nblk(n, 128)), dim3(128), 0, st, oa);
            } else {
                hipLaunchKernelGGL(k_outer<false>, dim3(nblk(n, 128)), dim3(128), 0, st, oa);
            }
            int fl[2] = {0, 0};
            HIPCHK(hipMemcpyAsync(fl, h->flags + 5, sizeof(fl), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            const int maxc = fl[0];
            need = fl[1];
            if (maxc <= h->ko) break;
            if (h->batch) {
                // a batch keeps its candidate capacity (growing it reallocates
                // [KO][LD] for every scene): a scene with a sphere beyond it is
                // failed like a blow-up, and the rebuild starts over without it
                std::vector<int> ovf(h->scap);
                HIPCHK(hipMemcpy(ovf.data(), h->ovf, sizeof(int) * h->scap, hipMemcpyDeviceToHost));
                int marked = 0;
                for (int k = 0; k < h->scap; ++k)
                    if (ovf[k] && h->sc[k].state == VG_SCENE_RUNNING) {
                        h->sc[k].state = VG_SCENE_RETIRING;
                        h->sc[k].failed = true;
                        ++marked;
                        if (getenv("VG_DEBUG")) fprintf(stderr, "[vg] scene %d failed: candidates > %d\n", k, h->ko);
                    }
                if (marked) return rebuild(h);
            }
            // grow geometrically: every growth reallocates [KO][LD] and reruns the pass
            const int ko_new = std::max(maxc + maxc / 2 + 8, 2 * h->ko);
            if (getenv("VG_DEBUG")) fprintf(stderr, "[vg] candidate capacity %d -> %d\n", h->ko, ko_new);
            if (maxc > 8192)
                return fail(h, VG_ECAPACITY, "a sphere has more than 8192 candidates (a scene is blowing up)");
            HIPCHK(hipFree(h->onbr));
            h->onbr = nullptr;
            h->ko = ko_new;
            if (dalloc(h, &h->onbr, (size_t)h->ko * h->ld)) return VG_ENOMEM;
            oa.nbr = h->onbr;
            oa.ko = h->ko;
        }
        pc.mark("candidates (synced)");
        // neighbour-list capacity: a sphere with more entries within the skin
        // than the lists hold grows both list buffers (up to KI_MAX entries),
        // with a margin for the lists' growth until the next rebuild; the new
        // lists are written at the new capacity, the previous ones are read
        // at theirs
        const int ki_old = h->ki;
        const bool grow = need > h->ki;
        if (grow) {
            if (need > KI_MAX)
                return fail(h, VG_ECAPACITY, "a sphere has more than " + std::to_string(KI_MAX) +
                                                 " neighbours within the skin of contact");
            const int ki_new = std::min(KI_MAX, need + need / 4 + 4);
            if (getenv("VG_DEBUG")) fprintf(stderr, "[vg] neighbour capacity %d -> %d\n", h->ki, ki_new);
            Ell &en = h->ell[h->ecur ^ 1];
            free_ell(en);
            if (alloc_ell(h, en, h->ld, ki_new)) return VG_ENOMEM;
            if (h->record) {
                // per-entry records of record steps: none are read across a rebuild
                if (h->rec) HIPCHK(hipFree(h->rec));
                h->rec = nullptr;
                if (dalloc(h, &h->rec, (size_t)ki_new * 13 * h->ld)) return VG_ENOMEM;
            }
            h->ki = ki_new;
        }
        // neighbour lists with history carry
        InnerArgs ia;
        ia.g = G1w; ia.gold = G0; ia.ell = h->ell[h->ecur ^ 1]; ia.eold = h->ell[h->ecur];
        ia.n = n; ia.ld = h->ld; ia.ki = h->ki; ia.ki_old = ki_old; ia.perm = h->perm; ia.cnt = h->ocount;
        ia.nbr = h->onbr; ia.ko = h->ko; ia.scenes = h->scenes; ia.models = h->models; ia.walls = h->walls;
        ia.flags = h->flags;
        // the previous lists' partners of each thread's sphere are staged in LDS
        const int ib = ki_old > 64 ? 64 : 128;
        hipLaunchKernelGGL(k_inner_build, dim3(nblk(n, ib)), dim3(ib), (size_t)ki_old * ib * sizeof(unsigned), st,
                           ia);
        if (h->any_constrained) {
            const int rc = build_cw_pairs(h, G1w, n);
            if (rc) return rc;
        }
        // zone of interest: the re-keyed records over the carried lists (SimMediator::Initiate)
        for (const ZoiPlan &P : zplans) {
            if (P.nrec == 0) continue;
            const SceneHost &S = h->sc[P.slot];
            unsigned *d_sid = nullptr, *d_wid = nullptr;
            if (id_maps(h, P.slot, &d_sid, &d_wid)) return VG_ENOMEM;
            const int b = h->sdev[P.slot].rank_off;
            hipLaunchKernelGGL(k_restore_history, dim3(nblk(S.n, 128)), dim3(128), 0, st, G1w, h->ell[h->ecur ^ 1],
                               b, b + S.n, h->ld, h->ki, (const unsigned *)d_sid, (const unsigned *)d_wid,
                               (const unsigned long long *)P.keys, (const RestoreRec *)P.recs, P.nrec,
                               (const ModelDev *)h->models, (const SceneDev *)h->scenes, (const LawDev *)h->laws,
                               (const WallDev *)h->walls);
            HIPCHK(hipStreamSynchronize(st));
            (void)hipFree(d_sid);
            (void)hipFree(d_wid);
        }
        if (grow) {
            // the previous lists, read above, become the spare buffer at the new capacity
            HIPCHK(hipStreamSynchronize(st));
            free_ell(h->ell[h->ecur]);
            if (alloc_ell(h, h->ell[h->ecur], h->ld, h->ki)) return VG_ENOMEM;
        }
    }
    h->n_iact = -1;   // counted on request (vg_num_interactions)
    // the fresh lists need no relist: clear the ring slot the next step reads
    // and restart the wall displacement bounds
    hipLaunchKernelGGL(k_ring_clear, dim3(nblk(h->scap, 256)), dim3(256), 0, st, h->rflags, h->scap,
                       (int)(h->step_count % 3));
    if (h->nw_used) hipLaunchKernelGGL(k_wall_path_reset, dim3(nblk(h->nw_used, 64)), dim3(64), 0, st, h->walls, h->nw_used);
    HIPCHK(hipGetLastError());
    // new layout: live scenes in slot order
    for (int k : live) {
        SceneHost &S = h->sc[k];
        S.seg_begin = h->sdev[k].rank_off;
        S.seg_end = S.seg_begin + S.n;
    }
    if (retiring) {
        for (int k = 0; k < h->scap; ++k) {
            SceneHost &S = h->sc[k];
            if (S.state != VG_SCENE_RETIRING) continue;
            for (int w = 0; w < S.nw; ++w) {
                h->wall_used[S.w_begin + w] = 0;
                h->walls_h[S.w_begin + w].scene = -1;
            }
            if (S.nw)
                HIPCHK(hipMemcpy(h->walls + S.w_begin, h->walls_h.data() + S.w_begin, sizeof(WallDev) * S.nw,
                                 hipMemcpyHostToDevice));
            S.state = S.failed ? VG_SCENE_FAILED : (S.result.empty() ? VG_SCENE_FREE : VG_SCENE_DONE);
            S.seg_begin = S.seg_end = 0;
            h->sdev[k] = SceneDev();
        }
        if (upload_scenes(h)) return VG_EHIP;
        update_batch_flags(h);
    }
    if (n > 0) {
        std::swap(h->geo[0].oid, h->oid_spare);
        std::swap(h->geo[0].mdl, h->mdl_spare);
        h->geo[1].oid = h->geo[0].oid;
        h->geo[1].mdl = h->geo[0].mdl;
    }
    h->n = n;
    h->cur ^= 1;
    h->ecur ^= 1;
    h->ent_valid = false;
    h->rebuilds++;
    pc.mark("inner lists launch, done");
    return 0;
}

// ---------------------------------------------------------------------------
// Allocation
// ---------------------------------------------------------------------------

static void set_err_create(const std::string &s) { g_create_error = s; }

extern "C" const char *vg_create_error(void) { return g_create_error.c_str(); }

extern "C" const char *vg_last_error(const vg_scene *h) { return h ? h->err.c_str() : g_create_error.c_str(); }

static int alloc_geo(vg_scene *h, Geo &g, int n)
{
    if (dalloc(h, &g.pos, n) || dalloc(h, &g.vel, n) || dalloc(h, &g.avl, n) || dalloc(h, &g.ori, 6 * (size_t)n))
        return VG_ENOMEM;
    return 0;
}

static void free_ell(Ell &e)
{
    for (void *p : {(void *)e.hdr, (void *)e.ent, (void *)e.slot, (void *)e.hd})
        if (p) (void)hipFree(p);
    e = Ell();
}

static int alloc_ell(vg_scene *h, Ell &e, int n, int ki)
{
    if (dalloc(h, &e.hdr, n) || dalloc(h, &e.ent, (size_t)2 * ki * n) || dalloc(h, &e.slot, (size_t)2 * ki * n) ||
        dalloc(h, &e.hd, (size_t)ki * n))
        return VG_ENOMEM;
    // the spring-dashpot kernels write the fourth history word as zero
    HIPCHK(hipMemset(e.hd, 0, sizeof(double4) * ki * (size_t)n));
    return 0;
}

static void free_scene(vg_scene *h)
{
    auto F = [](void *p) {
        if (p) (void)hipFree(p);
    };
    F(h->geo[0].oid); F(h->geo[0].mdl); F(h->oid_spare); F(h->mdl_spare);
    for (int b = 0; b < 2; ++b) {
        F(h->geo[b].pos); F(h->geo[b].vel); F(h->geo[b].avl); F(h->geo[b].ori);
        F(h->ell[b].hdr); F(h->ell[b].ent); F(h->ell[b].slot); F(h->ell[b].hd);
        F(h->xs[b]); F(h->ys[b]); F(h->zs[b]);
    }
    F(h->models); F(h->mm); F(h->laws); F(h->scenes); F(h->stime); F(h->rflags); F(h->walls); F(h->walls_ref);
    F(h->wkin); F(h->holes); F(h->pbc); F(h->assigned);
    F(h->pbc_ckey[0]); F(h->pbc_ckey[1]); F(h->pbc_cfirst[0]); F(h->pbc_cfirst[1]); F(h->pbc_ncell);
    F(h->pbc_pcount); F(h->pbc_poff); F(h->pbc_first); F(h->pbc_pairs); F(h->pbc_ekey[0]); F(h->pbc_ekey[1]);
    F(h->pbc_eval[0]); F(h->pbc_eval[1]); F(h->assigned_sph); F(h->en_part); F(h->en_out); F(h->flags); F(h->relists);
    F(h->contacts); F(h->red); F(h->inv); F(h->perm); F(h->svals); F(h->keys); F(h->skeys); F(h->qidx); F(h->segs);
    F(h->qval); F(h->qsel_tg); F(h->qsel_hist); F(h->qmap); F(h->safe); F(h->lin); F(h->cstart); F(h->cend); F(h->ocount); F(h->onbr); F(h->tmp); F(h->rec);
    F(h->rec_imp); F(h->rec_imp2); F(h->snap); F(h->snap_wimp); F(h->sid[0]); F(h->sid[1]); F(h->fetch); F(h->ovf);
    F(h->cw_val); F(h->cw_stamp); F(h->cw_key[0]); F(h->cw_key[1]); F(h->cw_vidx[0]); F(h->cw_vidx[1]);
    F(h->cw_w); F(h->cw_p); F(h->cw_cnt); F(h->cw_beg); F(h->cw_end); F(h->cw_imp);
    F(h->zoi_dead); F(h->zoi_cnt); F(h->zoi_rank); F(h->zoi_pre);
    if (h->pin) (void)hipHostFree(h->pin);
    for (auto e : h->kev) (void)hipEventDestroy(e);
    for (StepGraph &g : h->graphs) graph_free(g);
    for (auto e : h->rev) (void)hipEventDestroy(e);
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

// Scratch of vg_scene_packing_density_noalloc: the sample points (x, y, z
// each) followed by the 64-bit hit counter.  Every GPU the library runs on
// holds its own copy until the process ends.  In host code `scratch` is only
// the symbol's host-side stand-in: kernels and copies use the device address
// that scratch_setup looks up for each device.
constexpr int SCRATCH_POINTS = VG_DENSITY_NOALLOC_MAX_POINTS;
static __device__ __align__(256) double scratch[3 * SCRATCH_POINTS + 32];
constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[SCRATCH_MAX_DEVICES];

// Once per device, with that device current (handle set-up).
static int scratch_setup(vg_scene *h)
{
    if (h->device < 0 || h->device >= SCRATCH_MAX_DEVICES) return fail(h, VG_EUNSUPPORTED, "device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[h->device]) return 0;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
        return fail(h, VG_EHIP, "hipGetSymbolAddress of the scratch array failed");
    g_scratch_dev[h->device] = static_cast<double *>(p);
    return 0;
}

static double *scratch_of(int device)
{
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_dev[device];
}

// Device buffers for ld sphere slots, wcap wall slots and scap scene slots.
static int alloc_batch(vg_scene *h, int ld, int wcap, int scap, int ki, bool record)
{
    h->ld = ld; h->wcap = wcap; h->scap = scap; h->ki = ki; h->record = record;
    h->sc.assign(scap, SceneHost());
    h->sdev.assign(scap, SceneDev());
    h->walls_h.assign(std::max(wcap, 1), WallDev());
    h->wall_used.assign(std::max(wcap, 1), 0);
    if (hipGetDevice(&h->device) != hipSuccess) return fail(h, VG_EHIP, "no HIP device");
    if (int rc = scratch_setup(h)) return rc;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(h, VG_EHIP, "hipStreamCreate failed");
    (void)hipEventCreate(&h->ev0);
    (void)hipEventCreate(&h->ev1);
    int rc = 0;
    rc |= alloc_geo(h, h->geo[0], ld) | alloc_geo(h, h->geo[1], ld);
    // rank and model of a slot change only at rebuilds: both state buffers
    // share one copy, the rebuild's permutation writes the spare
    rc |= dalloc(h, &h->geo[0].oid, ld) | dalloc(h, &h->geo[0].mdl, ld);
    rc |= dalloc(h, &h->oid_spare, ld) | dalloc(h, &h->mdl_spare, ld);
    h->geo[1].oid = h->geo[0].oid;
    h->geo[1].mdl = h->geo[0].mdl;
    rc |= alloc_ell(h, h->ell[0], ld, ki) | alloc_ell(h, h->ell[1], ld, ki);
    rc |= dalloc(h, &h->scenes, scap) | dalloc(h, &h->stime, 2 * (size_t)scap) | dalloc(h, &h->rflags, 3 * (size_t)scap);
    rc |= dalloc(h, &h->walls, wcap) | dalloc(h, &h->walls_ref, wcap) | dalloc(h, &h->wkin, wcap);
    rc |= dalloc(h, &h->assigned, (size_t)15 * wcap);
    rc |= dalloc(h, &h->cw_beg, std::max(wcap, 1)) | dalloc(h, &h->cw_end, std::max(wcap, 1)) |
          dalloc(h, &h->cw_imp, (size_t)3 * std::max(wcap, 1));
    rc |= dalloc(h, &h->flags, 8) | dalloc(h, &h->relists, 1) | dalloc(h, &h->red, (size_t)8 * scap);
    rc |= dalloc(h, &h->inv, ld) | dalloc(h, &h->perm, ld) | dalloc(h, &h->svals, ld);
    rc |= dalloc(h, &h->keys, ld) | dalloc(h, &h->skeys, ld);
    for (int b = 0; b < 2; ++b) rc |= dalloc(h, &h->xs[b], ld) | dalloc(h, &h->ys[b], ld) | dalloc(h, &h->zs[b], ld);
    rc |= dalloc(h, &h->qidx, 2 * (size_t)scap) | dalloc(h, &h->segs, 2 * (size_t)scap) | dalloc(h, &h->qval, 6 * (size_t)scap);
    rc |= dalloc(h, &h->qsel_tg, 6 * (size_t)scap) | dalloc(h, &h->qsel_hist, (size_t)6 * 256 * scap) |
          dalloc(h, &h->qmap, (size_t)scap);
    rc |= dalloc(h, &h->sid[0], ld) | dalloc(h, &h->sid[1], ld) | dalloc(h, &h->ovf, scap);
    rc |= dalloc(h, &h->safe, ld) | dalloc(h, &h->lin, ld);
    rc |= dalloc(h, &h->ocount, (size_t)ld + 1);
    rc |= dalloc(h, &h->rec_imp, (size_t)6 * ld) | dalloc(h, &h->rec_imp2, (size_t)6 * ld);
    if (record) rc |= dalloc(h, &h->rec, (size_t)ki * 13 * ld) | dalloc(h, &h->snap_wimp, (size_t)6 * std::max(wcap, 1));
    rc |= dalloc(h, &h->en_out, 24) | dalloc(h, &h->en_part, (size_t)16 * 1024);
    rc |= dalloc(h, &h->contacts, nblk(ld, 64) + 1);
    if (rc) return VG_ENOMEM;
    if (dalloc(h, &h->onbr, (size_t)h->ko * ld)) return VG_ENOMEM;
    if (hipMemset(h->flags, 0, 8 * sizeof(int)) != hipSuccess ||
        hipMemset(h->relists, 0, sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(h->contacts, 0, sizeof(unsigned long long) * (nblk(ld, 64) + 1)) != hipSuccess ||
        hipMemset(h->rflags, 0, sizeof(int) * 3 * scap) != hipSuccess ||
        hipMemset(h->scenes, 0, sizeof(SceneDev) * scap) != hipSuccess)
        return fail(h, VG_EHIP, "device initialisation failed");
    // wall slots not in use belong to no scene (the wall kernel skips them)
    for (WallDev &W : h->walls_h) W.scene = -1;
    if (hipMemcpy(h->walls, h->walls_h.data(), sizeof(WallDev) * wcap, hipMemcpyHostToDevice) != hipSuccess)
        return fail(h, VG_EHIP, "device initialisation failed");
    return 0;
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
    SceneHost probe;
    std::string err;
    int rc = scene_from_desc(h, d, probe, err);
    if (rc) {
        h->err = err;
        return bail(rc);
    }
    int ki = 48;
    const char *ki_env = getenv("VG_NEIGHBOUR_CAPACITY");
    if (ki_env) ki = std::max(8, std::min(KI_MAX, atoi(ki_env)));   // a start: rebuilds grow it as needed
    h->up_iact = d->up_iact ? d->up_iact : 1;
    rc = alloc_batch(h, probe.n, std::max(probe.nw, 1), 1, ki, true);
    if (rc) return bail(rc);
    int slot = -1;
    rc = admit(h, d, -1, &slot);
    if (rc) return bail(rc);
    // Consultant::Reset + SimMediator::Initiate at creation
    if (rebuild(h)) return bail(VG_EHIP);
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
        h->err = "device error during the first candidate-list build";
        return bail(VG_EHIP);
    }
    *out = h;
    return 0;
}

extern "C" int vg_batch_create(const vg_batch_desc *bd, vg_scene **out)
{
    if (!bd || !out || bd->max_spheres <= 0 || bd->max_scenes <= 0 || bd->max_walls < 0) {
        set_err_create("invalid batch description");
        return VG_EINVAL;
    }
    *out = nullptr;
    if (bd->max_spheres + (long long)bd->max_walls >= (long long)PARTNER_MASK) {
        set_err_create("too many DOs for the neighbour-list encoding");
        return VG_EUNSUPPORTED;
    }
    vg_scene *h = new vg_scene();
    h->batch = true;
    h->up_iact = bd->up_iact ? bd->up_iact : 1;
    // VG_GRAPHS=0: every step launched from the host (for tools that trace
    // kernel launches and do not follow graph replays)
    if (const char *g = getenv("VG_GRAPHS")) h->use_graphs = atoi(g) != 0;
    // large layouts: start the candidate capacity high, since growing
    // reallocates [KO][LD] and stalls the device for seconds at 1e7 slots
    h->ko = 512;
    const int ki = bd->neighbour_capacity > 0 ? std::max(8, std::min(KI_MAX, bd->neighbour_capacity)) : 32;
    int rc = alloc_batch(h, (int)bd->max_spheres, std::max(bd->max_walls, 1), bd->max_scenes, ki, false);
    if (rc) {
        set_err_create(h->err);
        free_scene(h);
        delete h;
        return rc;
    }
    *out = h;
    return 0;
}

extern "C" int vg_admit(vg_scene *h, const vg_scene_desc *d, long long steps, int *scene)
{
    if (!h || !d || !scene) return VG_EINVAL;
    if (!h->batch) return fail(h, VG_EINVAL, "vg_admit needs a batch handle (vg_batch_create)");
    HIPCHK(hipStreamSynchronize(h->stream));
    return admit(h, d, steps, scene);
}

// ---------------------------------------------------------------------------
// Stepping
// ---------------------------------------------------------------------------

// one scene slot whose spheres share one group: the step reads its scene row
// and sphere-sphere law through uniform addresses
static bool uniform_scene(const vg_scene *h)
{
    return h->scap == 1 && h->single_group && !h->sc.empty() && h->sc[0].state == VG_SCENE_RUNNING;
}

static StepArgs step_args(vg_scene *h)
{
    StepArgs a;
    a.n = h->n; a.ld = h->ld; a.nw = h->nw_used; a.ki = h->ki;
    a.in = h->geo[h->cur]; a.out = h->geo[h->cur ^ 1];
    a.ell = h->ell[h->ecur];
    a.o_cnt = h->ocount; a.o_nbr = h->onbr; a.ko = h->ko;
    a.t = tables(h);
    a.scenes = h->scenes;
    a.n_scenes = h->scap;
    a.walls = h->walls;
    a.rflags = h->rflags; a.flags = h->flags; a.relists = h->relists; a.contacts = h->contacts;
    a.step = (int)(h->step_count % 3);
    a.rec_imp = h->rec_imp; a.rec = h->rec;
    a.holes = h->holes;
    a.pbc = h->pbc;
    a.single_group = h->single_group;
    a.no_bond = h->no_bond;
    a.group0 = 0;
    a.law_ss = 0;
    a.any_fext = 0;
    for (const ModelDev &m : h->models_h)
        if (m.fext[0] != 0.0 || m.fext[1] != 0.0 || m.fext[2] != 0.0) a.any_fext = 1;
    if (uniform_scene(h)) {
        const SceneHost &S = h->sc[0];
        a.group0 = S.group0;
        a.law_ss = S.law_off + S.group0 * S.n_groups + S.group0;
    }
    a.cw_w = h->cw_w; a.cw_p = h->cw_p; a.cw_val = h->cw_val; a.cw_stamp = h->cw_stamp;
    a.stamp = (unsigned)(h->step_count & 0x7fffffffull);
    return a;
}

static const SceneHost *scene0(const vg_scene *h)
{
    return (!h || h->batch || h->sc.empty()) ? nullptr : &h->sc[0];
}

static int export_active(vg_scene *h, int slot, ActiveOut *out, size_t cap, unsigned long long *dcnt,
                         unsigned *d_sid, unsigned *d_wid, bool backup)
{
    const SceneHost &S = h->sc[slot];
    HIPCHK(hipMemsetAsync(dcnt, 0, sizeof(unsigned long long), h->stream));
    hipLaunchKernelGGL((h->holes || h->pbc) ? k_export_active<true> : k_export_active<false>, dim3(nblk(S.n, 256)),
                       dim3(256),
                       0, h->stream, h->geo[h->cur], h->ell[h->ecur], S.seg_begin, S.seg_end, h->ld, h->ki,
                       (const unsigned *)h->ocount, (const unsigned *)h->onbr, (const WallDev *)h->walls,
                       (const unsigned *)d_wid, (const unsigned *)d_sid,
                       h->ent_valid ? (const double *)h->rec : (const double *)nullptr, out, (unsigned long long)cap,
                       dcnt, backup ? 1 : 0, h->no_bond, (const double4 *)h->holes, (const PbcDev *)h->pbc);
    HIPCHK(hipGetLastError());
    return 0;
}

// DO-id maps of a scene on the device (for exports): sphere rank -> DO id,
// wall slot -> DO id
static int id_maps(vg_scene *h, int slot, unsigned **d_sid, unsigned **d_wid)
{
    const SceneHost &S = h->sc[slot];
    std::vector<unsigned> wid(h->wcap, 0u);
    for (int w = 0; w < S.nw; ++w) wid[S.w_begin + w] = S.wall_ids[w];
    if (dalloc(h, d_sid, (size_t)S.n) || dalloc(h, d_wid, (size_t)h->wcap)) return VG_ENOMEM;
    HIPCHK(hipMemcpy(*d_sid, S.sphere_ids.data(), sizeof(unsigned) * S.n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(*d_wid, wid.data(), sizeof(unsigned) * h->wcap, hipMemcpyHostToDevice));
    return 0;
}

// RebuildIactRecordTab before the candidate-list reset (NBSConsultant.cpp:40-47):
// the active contacts and the wall impulses of the last step are kept, since
// the rebuild re-sorts the spheres and their lists (one-scene handles).
static int snapshot_records(vg_scene *h)
{
    unsigned long long *dcnt = nullptr;
    unsigned *d_sid = nullptr, *d_wid = nullptr;
    if (dalloc(h, &dcnt, 1) || id_maps(h, 0, &d_sid, &d_wid)) return VG_ENOMEM;
    if (export_active(h, 0, nullptr, 0, dcnt, d_sid, d_wid, false)) return VG_EHIP;
    unsigned long long cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, dcnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (cnt > h->snap_cap) {
        if (h->snap) HIPCHK(hipFree(h->snap));
        h->snap = nullptr;
        h->snap_cap = cnt + cnt / 4 + 64;
        if (dalloc(h, &h->snap, h->snap_cap)) return VG_ENOMEM;
    }
    if (export_active(h, 0, h->snap, h->snap_cap, dcnt, d_sid, d_wid, false)) return VG_EHIP;
    const SceneHost &S = h->sc[0];
    if (S.nw && h->ent_valid)
        hipLaunchKernelGGL(k_wall_impulse, dim3(S.nw), dim3(256), 0, h->stream, h->ell[h->ecur], S.seg_begin,
                           S.seg_end, h->ld, h->ki, S.w_begin, (const double *)h->rec, h->no_bond, h->snap_wimp);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(dcnt));
    HIPCHK(hipFree(d_sid));
    HIPCHK(hipFree(d_wid));
    h->snap_count = (long long)cnt;
    h->snap_valid = true;
    return 0;
}

// LeapConsultant::NextStep's rebuild (candidate list, history carry, zone of
// interest), bracketed by timing events while kernel timing is on.
static int do_rebuild(vg_scene *h)
{
    const bool rt = 2 * (h->rev_used + 1) <= (int)h->rev.size();
    if (rt) HIPCHK(hipEventRecord(h->rev[2 * h->rev_used], h->stream));
    const int rc = rebuild(h);
    if (rt && !rc) {
        HIPCHK(hipEventRecord(h->rev[2 * h->rev_used + 1], h->stream));
        h->rev_used++;
    }
    return rc;
}

// A deferred rebuild runs before anything that needs the rebuilt lists: the
// next step (the same device work in the same order as a rebuild at the end of
// the previous step), or a query of the candidate list.  With keep_records the
// step's records are kept for the exports first, as an immediate rebuild does.
static int flush_rebuild(vg_scene *h, bool keep_records)
{
    if (!h->rebuild_pending) return 0;
    h->rebuild_pending = false;
    if (keep_records && h->ent_valid && snapshot_records(h)) return VG_EHIP;
    return do_rebuild(h);
}

// The step kernel's law set for the placed scenes: spring-dashpot only,
// ISwH only, all five, or all five with the extended geometry.
static int law_set(const vg_scene *h)
{
    return (h->holes || h->pbc) ? LS_EXT : (h->sd_only ? LS_SD : (h->h_only ? LS_H : LS_ALL));
}

static int launch_step(vg_scene *h, bool rec, bool kt);
static int after_step(vg_scene *h, bool rec, const std::vector<std::pair<int, int>> *assign_spheres);

static int one_step(vg_scene *h, bool rec, const std::vector<std::pair<int, int>> *assign_spheres)
{
    if (int rc = flush_rebuild(h, false)) return rc;
    if (h->n == 0) {
        // nothing placed yet: the first boundary places the queued scenes
        bool queued = false;
        for (const SceneHost &S : h->sc) queued = queued || S.state == VG_SCENE_QUEUED;
        if (!queued) return 0;
        if (rebuild(h)) return VG_EHIP;
        if (h->n == 0) return 0;
    }
    rec = rec && h->record;
    const bool kt = 2 * (h->kev_used + 1) <= (int)h->kev.size();
    if (kt) HIPCHK(hipEventRecord(h->kev[2 * h->kev_used], h->stream));
    if (int rc = launch_step(h, rec, kt)) return rc;
    h->cur ^= 1;
    h->step_count++;
    h->round++;
    return after_step(h, rec, assign_spheres);
}

// The kernels of one step on the handle's stream (k_step, the constrained
// walls' impulse sums, the wall responses), for the handle's current buffers
// and step count.
static int launch_step(vg_scene *h, bool rec, bool kt)
{
    StepArgs a = step_args(h);
    const dim3 grid(nblk(std::max(h->n, h->scap), STEP_NT)), blk(STEP_NT);
    // kernel specialisation: law set of the placed scenes, record path, precision
    const int ls = law_set(h);
    const bool mix = h->precision == VG_PRECISION_MIXED;
    using KFn = void (*)(StepArgs);
    static const KFn kfn[2][2][4][2] = {
        {{{k_step<false, LS_SD, false, false>, k_step<false, LS_SD, true, false>},
          {k_step<false, LS_H, false, false>, k_step<false, LS_H, true, false>},
          {k_step<false, LS_ALL, false, false>, k_step<false, LS_ALL, true, false>},
          {k_step<false, LS_EXT, false, false>, k_step<false, LS_EXT, true, false>}},
         {{k_step<true, LS_SD, false, false>, k_step<true, LS_SD, true, false>},
          {k_step<true, LS_H, false, false>, k_step<true, LS_H, true, false>},
          {k_step<true, LS_ALL, false, false>, k_step<true, LS_ALL, true, false>},
          {k_step<true, LS_EXT, false, false>, k_step<true, LS_EXT, true, false>}}},
        {{{k_step<false, LS_SD, false, true>, k_step<false, LS_SD, true, true>},
          {k_step<false, LS_H, false, true>, k_step<false, LS_H, true, true>},
          {k_step<false, LS_ALL, false, true>, k_step<false, LS_ALL, true, true>},
          {k_step<false, LS_EXT, false, false>, k_step<false, LS_EXT, true, false>}},
         {{k_step<true, LS_SD, false, true>, k_step<true, LS_SD, true, true>},
          {k_step<true, LS_H, false, true>, k_step<true, LS_H, true, true>},
          {k_step<true, LS_ALL, false, true>, k_step<true, LS_ALL, true, true>},
          {k_step<true, LS_EXT, false, false>, k_step<true, LS_EXT, true, false>}}}};
    // lists longer than 64 entries: the wide variants, for the general law sets
    static const KFn kfn_wide[2][2][2][2] = {
        {{{k_step<false, LS_ALL, false, false, 2>, k_step<false, LS_ALL, true, false, 2>},
          {k_step<false, LS_EXT, false, false, 2>, k_step<false, LS_EXT, true, false, 2>}},
         {{k_step<true, LS_ALL, false, false, 2>, k_step<true, LS_ALL, true, false, 2>},
          {k_step<true, LS_EXT, false, false, 2>, k_step<true, LS_EXT, true, false, 2>}}},
        {{{k_step<false, LS_ALL, false, true, 2>, k_step<false, LS_ALL, true, true, 2>},
          {k_step<false, LS_EXT, false, false, 2>, k_step<false, LS_EXT, true, false, 2>}},
         {{k_step<true, LS_ALL, false, true, 2>, k_step<true, LS_ALL, true, true, 2>},
          {k_step<true, LS_EXT, false, false, 2>, k_step<true, LS_EXT, true, false, 2>}}}};
    const KFn fn = h->ki > 64 ? kfn_wide[uniform_scene(h) ? 1 : 0][rec ? 1 : 0][ls == LS_EXT ? 1 : 0][mix ? 1 : 0]
                              : kfn[uniform_scene(h) ? 1 : 0][rec ? 1 : 0][ls][mix ? 1 : 0];
    hipLaunchKernelGGL(fn, grid, blk, 0, h->stream, a);
    if (kt) {
        HIPCHK(hipEventRecord(h->kev[2 * h->kev_used + 1], h->stream));
        h->kev_used++;
    }
    h->last_launches++;
    if (h->nw_used && h->any_constrained)
        hipLaunchKernelGGL(k_cw_sum, dim3(nblk(h->nw_used, 64)), dim3(64), 0, h->stream, (const WallDev *)h->walls,
                           h->nw_used, (const SceneDev *)h->scenes, (const int *)h->cw_beg, (const int *)h->cw_end,
                           (const double *)h->cw_val, (const unsigned *)h->cw_stamp, a.stamp, h->cw_imp);
    if (h->nw_used)
        hipLaunchKernelGGL(k_walls, dim3(nblk(h->nw_used, 64)), dim3(64), 0, h->stream, h->walls, h->nw_used,
                           (const SceneDev *)h->scenes, h->stime, (const double *)h->assigned,
                           (const WallDev *)h->walls_ref, (const WallKin *)h->wkin, (const double *)h->cw_imp,
                           h->rflags, (int)(h->step_count % 6));
    return 0;
}

// The host side of a step after its launches: scene times and step counts,
// scenes that reach their step count, assigned sphere statuses, and the
// rebuild at the end of a period.
static int after_step(vg_scene *h, bool rec, const std::vector<std::pair<int, int>> *assign_spheres)
{
    // SystemParameter::NextStep per scene; a scene that reaches its step
    // count is frozen (its state is held until the next boundary)
    bool frozen = false;
    for (int k = 0; k < h->scap; ++k) {
        SceneHost &S = h->sc[k];
        if (S.state != VG_SCENE_RUNNING) continue;
        S.t += S.dt;
        S.steps_done++;
        if (S.steps_target >= 0 && S.steps_done >= S.steps_target) {
            S.state = VG_SCENE_FINISHED;
            h->sdev[k].active = 0;
            frozen = true;
        }
    }
    if (frozen && upload_scenes(h)) return VG_EHIP;
    if (assign_spheres) {
        // SimMediator::Run(map): the assigned statuses replace the integrated
        // ones before contact detection and the rebuild; positions may jump,
        // so the scene's spheres relist at the next step
        for (const auto &pr : *assign_spheres) {
            const SceneHost &S = h->sc[pr.first];
            hipLaunchKernelGGL(k_assign_spheres, dim3(nblk(S.n, 256)), dim3(256), 0, h->stream, h->geo[h->cur],
                               S.seg_begin, S.seg_end, h->ld, (const double *)h->assigned_sph + pr.second,
                               rec ? h->rec_imp : (double *)nullptr);
            hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, h->stream,
                               h->rflags + 3 * pr.first + (h->step_count % 3), 1ll, 1);
        }
    }
    h->rec_valid = rec;
    h->ent_valid = rec;
    h->snap_valid = false;
    HIPCHK(hipGetLastError());
    if (h->round % h->up_iact == 0) {
        // the last step of a call: its records are read in place until the
        // handle is next used (a zone-of-interest clean-up erases DOs at the
        // rebuild, which the exports must see: that rebuild runs now)
        if (rec && !h->batch && !h->any_zoi) {
            h->rebuild_pending = true;
            return 0;
        }
        if (rec && !h->batch && snapshot_records(h)) return VG_EHIP;
        return do_rebuild(h);
    }
    return 0;
}

// A batch handle steps a whole rebuild period as one HIP graph when nothing
// between its steps needs the host: the period starts at a rebuild, no
// scene reaches its step count inside it, no wall is constrained (their
// per-step stamps), and kernel timing is off.  The launches of the period
// are captured once per set of launch arguments (buffers, counts, the step
// count modulo 6) and replayed; the host's per-step bookkeeping and the
// rebuild follow the replay as they follow eager launches.
static bool graph_period_ok(const vg_scene *h, int remaining)
{
    if (!h->batch || !h->use_graphs || h->n == 0 || h->any_constrained || h->rebuild_pending) return false;
    if (remaining < (int)h->up_iact || h->round % h->up_iact != 0) return false;
    if (2 * (h->kev_used + 1) <= (int)h->kev.size()) return false;
    for (const SceneHost &S : h->sc)
        if (S.state == VG_SCENE_RUNNING && S.steps_target >= 0 && S.steps_target - S.steps_done < (long long)h->up_iact)
            return false;
    return true;
}

static int graph_period(vg_scene *h)
{
    const unsigned U = h->up_iact;
    // the launch arguments of the period's first step determine all of them
    StepArgs a0 = step_args(h);
    a0.stamp = 0;
    // with the kernel specialisation launch_step chooses (law set, precision, one-scene path)
    const long long extra[7] = {(long long)(h->step_count % 6), h->nw_used, (long long)U, h->precision,
                                (long long)uniform_scene(h), law_set(h), h->ki};
    std::vector<unsigned char> key(sizeof a0 + sizeof extra);
    std::memcpy(key.data(), &a0, sizeof a0);
    std::memcpy(key.data() + sizeof a0, extra, sizeof extra);
    StepGraph *g = nullptr;
    for (StepGraph &c : h->graphs)
        if (c.exec && c.key == key) g = &c;
    if (!g) {
        // least recently used slot
        g = &h->graphs[0];
        for (StepGraph &c : h->graphs)
            if (!c.exec || c.last_use < g->last_use) g = &c;
        graph_free(*g);
        const int cur = h->cur;
        const unsigned long long step_count = h->step_count, round = h->round;
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        int rc = 0;
        for (unsigned k = 0; k < U && !rc; ++k) {
            rc = launch_step(h, false, false);
            h->cur ^= 1;
            h->step_count++;
            h->round++;
        }
        h->cur = cur;
        h->step_count = step_count;
        h->round = round;
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(h->stream, &graph);
        if (rc || e != hipSuccess || !graph) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc ? rc : fail(h, VG_EHIP, "capturing the step launches failed");
        }
        const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            g->exec = nullptr;
            return fail(h, VG_EHIP, "instantiating the step graph failed");
        }
        g->key = key;
    }
    g->last_use = ++h->graph_clock;
    HIPCHK(hipGraphLaunch(g->exec, h->stream));
    h->graph_replays++;
    h->last_launches += (int)U;
    for (unsigned k = 0; k < U; ++k) {
        h->cur ^= 1;
        h->step_count++;
        h->round++;
        if (int rc = after_step(h, false, nullptr)) return rc;
    }
    return 0;
}

extern "C" int vg_set_graphs(vg_scene *h, int on)
{
    if (!h) return VG_EINVAL;
    h->use_graphs = on != 0;
    if (!h->use_graphs)
        for (StepGraph &g : h->graphs) graph_free(g);
    return 0;
}

extern "C" int vg_graph_replays(vg_scene *h, long long *replays)
{
    if (!h || !replays) return VG_EINVAL;
    *replays = h->graph_replays;
    return 0;
}

extern "C" int vg_step(vg_scene *h, int nsteps)
{
    if (!h || nsteps < 0) return VG_EINVAL;
    h->last_launches = 0;
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    for (int k = 0; k < nsteps;) {
        if (graph_period_ok(h, nsteps - k)) {
            if (int rc = graph_period(h)) return rc;
            k += (int)h->up_iact;
            continue;
        }
        int rc = one_step(h, k == nsteps - 1, nullptr);
        if (rc) return rc;
        ++k;
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
        return fail(h, VG_ECAPACITY,
                    "a neighbour list outgrew its capacity between rebuilds (VG_NEIGHBOUR_CAPACITY or the batch's "
                    "neighbour_capacity sets a larger start, at most 127)");
    if (flags[3] & ERR_CW)
        return fail(h, VG_EUNSUPPORTED, "a sphere is a candidate of more than four constrained walls");
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
    const SceneHost *S0 = scene0(h);
    if (!S0) return fail(h, VG_EINVAL, "vg_step_assigned needs a one-scene handle");
    if (S0->has_pbc) return fail(h, VG_EUNSUPPORTED, "vg_step_assigned: not with periodic boundaries");
    const SceneHost &S = *S0;
    std::vector<double> buf((size_t)15 * h->wcap, 0.0);
    std::vector<char> set(h->wcap, 0);
    std::vector<double> sb;
    bool any_sphere = false;
    for (int k = 0; k < n; ++k) {
        if (ids[k] >= (unsigned)S.n_do) return fail(h, VG_EINVAL, "assigned status of an unknown DO id");
        const double *src[5] = {pos + 3 * k, vel + 3 * k, ox + 3 * k, oz + 3 * k, av + 3 * k};
        if (S.do_to_sphere[ids[k]] >= 0) {
            if (!any_sphere) sb.assign((size_t)16 * S.n, 0.0);
            any_sphere = true;
            double *b = &sb[(size_t)16 * S.do_to_sphere[ids[k]]];
            for (int f = 0; f < 5; ++f)
                for (int c = 0; c < 3; ++c) b[3 * f + c] = src[f][c];
            b[15] = 1.0;
            continue;
        }
        const int w = S.do_to_wall[ids[k]];
        double *b = &buf[(size_t)15 * w];
        for (int f = 0; f < 5; ++f)
            for (int c = 0; c < 3; ++c) b[3 * f + c] = src[f][c];
        set[w] = 1;
    }
    if (any_sphere) {
        if (!h->assigned_sph && dalloc(h, &h->assigned_sph, (size_t)16 * h->ld)) return VG_ENOMEM;
        HIPCHK(hipMemcpyAsync(h->assigned_sph, sb.data(), sizeof(double) * sb.size(), hipMemcpyHostToDevice,
                              h->stream));
    }
    bool changed = false;
    for (int w = S.w_begin; w < S.w_begin + S.nw; ++w) {
        const int want = set[w] ? 1 : (h->walls_h[w].kin == 1 ? 0 : h->walls_h[w].kin);
        if (want != h->walls_h[w].kin) {
            h->walls_h[w].kin = want;
            changed = true;
        }
    }
    if (S.nw) {
        HIPCHK(hipMemcpyAsync(h->assigned, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice, h->stream));
        if (changed) {
            // only the kin field changes here; the device copy holds the live status
            std::vector<WallDev> cur(h->wcap);
            HIPCHK(hipMemcpyAsync(cur.data(), h->walls, sizeof(WallDev) * h->wcap, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            for (int w = S.w_begin; w < S.w_begin + S.nw; ++w) cur[w].kin = h->walls_h[w].kin;
            HIPCHK(hipMemcpyAsync(h->walls, cur.data(), sizeof(WallDev) * h->wcap, hipMemcpyHostToDevice, h->stream));
        }
    }
    h->last_launches = 0;
    std::vector<std::pair<int, int>> asg;
    if (any_sphere) asg.push_back({0, 0});
    int rc = one_step(h, true, any_sphere ? &asg : nullptr);
    if (rc) return rc;
    // synchronous: the host buffers are reused by the next call
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

static int set_rotation(vg_scene *h, int slot, int n, const unsigned *ids, const double *c, const double *a,
                        double omega, double t_ref)
{
    const SceneHost &S = h->sc[slot];
    if (S.has_pbc) return fail(h, VG_EUNSUPPORTED, "wall rotation: not with periodic boundaries (walls stay fixed)");
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<WallDev> cur(h->wcap), ref(h->wcap);
    std::vector<WallKin> kin(h->wcap);
    HIPCHK(hipMemcpy(cur.data(), h->walls, sizeof(WallDev) * h->wcap, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ref.data(), h->walls_ref, sizeof(WallDev) * h->wcap, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(kin.data(), h->wkin, sizeof(WallKin) * h->wcap, hipMemcpyDeviceToHost));
    const double al = olen(a);
    for (int k = 0; k < n; ++k) {
        if (ids[k] >= (unsigned)S.n_do || S.do_to_wall[ids[k]] < 0)
            return fail(h, VG_EINVAL, "wall rotation applies to walls only");
        const int w = S.do_to_wall[ids[k]];
        kin[w].c = mk(c[0], c[1], c[2]);
        kin[w].a = mk(a[0] / al, a[1] / al, a[2] / al);
        kin[w].omega = omega;
        kin[w].t_ref = t_ref;
        cur[w].kin = 2;
        ref[w] = cur[w];
        h->walls_h[w].kin = 2;
    }
    HIPCHK(hipMemcpy(h->walls_ref, ref.data(), sizeof(WallDev) * h->wcap, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->walls, cur.data(), sizeof(WallDev) * h->wcap, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->wkin, kin.data(), sizeof(WallKin) * h->wcap, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int vg_set_wall_rotation(vg_scene *h, int n, const unsigned *ids, const double *c, const double *a,
                                    double omega, double t_ref)
{
    if (!h || n < 0) return VG_EINVAL;
    if (!scene0(h)) return fail(h, VG_EINVAL, "vg_set_wall_rotation needs a one-scene handle");
    return set_rotation(h, 0, n, ids, c, a, omega, t_ref);
}

// ---------------------------------------------------------------------------
// State and records
// ---------------------------------------------------------------------------

// DOStatus of every DO of a placed scene, [n_do][21]: pos vel Ox Oz w impact
// angular-impact, in the scene's DO-id order.
static int fetch_scene_state(vg_scene *h, int slot, std::vector<double> &hb)
{
    const SceneHost &S = h->sc[slot];
    HIPCHK(hipStreamSynchronize(h->stream));
    const int nd = S.n_do, n = S.n;
    if ((size_t)21 * n > h->fetch_cap) {
        if (h->fetch) HIPCHK(hipFree(h->fetch));
        h->fetch = nullptr;
        h->fetch_cap = (size_t)21 * n;
        if (dalloc(h, &h->fetch, h->fetch_cap)) return VG_ENOMEM;
    }
    double *dout = h->fetch;
    // pinned staging: a pageable download of a large scene runs several
    // times slower and holds the stream meanwhile
    if ((size_t)21 * n > h->pin_cap) {
        if (h->pin) HIPCHK(hipHostFree(h->pin));
        h->pin = nullptr;
        h->pin_cap = 0;
        HIPCHK(hipHostMalloc((void **)&h->pin, sizeof(double) * 21 * (size_t)std::max(n, 1), hipHostMallocDefault));
        h->pin_cap = (size_t)21 * std::max(n, 1);
    }
    double *sb = h->pin;
    hipLaunchKernelGGL(k_scatter_state, dim3(nblk(n, 256)), dim3(256), 0, h->stream, h->geo[h->cur], S.seg_begin,
                       S.seg_end, h->ld, h->rec_valid ? (const double *)h->rec_imp : (const double *)nullptr, dout);
    std::vector<double> wimp((size_t)6 * std::max(S.nw, 1), 0.0);
    double *dw = nullptr;
    const bool one = !h->batch;
    if (one && S.nw && h->snap_valid) {
        if (dalloc(h, &dw, (size_t)6 * S.nw)) return VG_ENOMEM;
        HIPCHK(hipMemcpyAsync(dw, h->snap_wimp, sizeof(double) * 6 * S.nw, hipMemcpyDeviceToDevice, h->stream));
    } else if (one && S.nw && h->ent_valid) {
        if (dalloc(h, &dw, (size_t)6 * S.nw)) return VG_ENOMEM;
        hipLaunchKernelGGL(k_wall_impulse, dim3(S.nw), dim3(256), 0, h->stream, h->ell[h->ecur], S.seg_begin,
                           S.seg_end, h->ld, h->ki, S.w_begin, (const double *)h->rec, h->no_bond, dw);
    }
    HIPCHK(hipMemcpyAsync(sb, dout, sizeof(double) * 21 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
    if (dw) HIPCHK(hipMemcpyAsync(wimp.data(), dw, sizeof(double) * 6 * S.nw, hipMemcpyDeviceToHost, h->stream));
    std::vector<WallDev> walls(std::max(S.nw, 1));
    if (S.nw)
        HIPCHK(hipMemcpyAsync(walls.data(), h->walls + S.w_begin, sizeof(WallDev) * S.nw, hipMemcpyDeviceToHost,
                              h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (dw) HIPCHK(hipFree(dw));
    hb.assign((size_t)21 * nd, 0.0);
    for (int s = 0; s < n; ++s)
        std::memcpy(&hb[(size_t)21 * S.sphere_ids[s]], &sb[(size_t)21 * s], 21 * sizeof(double));
    for (int w = 0; w < S.nw; ++w) {
        double *d = &hb[(size_t)21 * S.wall_ids[w]];
        const WallDev &W = walls[w];
        const d3 v[5] = {W.pos, W.vel, W.ox, W.oz, W.av};
        for (int k = 0; k < 5; ++k) { d[3 * k] = v[k].x; d[3 * k + 1] = v[k].y; d[3 * k + 2] = v[k].z; }
        // assigned statuses carry zero stored impulses (DOContainer::SetDOStatus)
        const bool zero = W.kin != 0;
        for (int c = 0; c < 6; ++c) d[15 + c] = zero ? 0.0 : wimp[(size_t)6 * w + c];
    }
    return 0;
}

static void split_state(const std::vector<double> &hb, int nd, double *pos, double *vel, double *ox, double *oz,
                        double *av, double *impact, double *ang_impact)
{
    double *outs[7] = {pos, vel, ox, oz, av, impact, ang_impact};
    for (int f = 0; f < 7; ++f) {
        if (!outs[f]) continue;
        for (int i = 0; i < nd; ++i)
            for (int c = 0; c < 3; ++c) outs[f][3 * i + c] = hb[(size_t)21 * i + 3 * f + c];
    }
}

extern "C" int vg_get_state(vg_scene *h, double *pos, double *vel, double *ox, double *oz, double *av,
                            double *impact, double *ang_impact)
{
    if (!h) return VG_EINVAL;
    const SceneHost *S = scene0(h);
    if (!S) return fail(h, VG_EINVAL, "vg_get_state needs a one-scene handle (vg_scene_get_state for batches)");
    std::vector<double> hb;
    if (fetch_scene_state(h, 0, hb)) return VG_EHIP;
    split_state(hb, S->n_do, pos, vel, ox, oz, av, impact, ang_impact);
    return 0;
}

extern "C" double vg_time(const vg_scene *h)
{
    const SceneHost *S = scene0(h);
    return S ? S->t : 0.0;
}

extern "C" int vg_num_interactions(vg_scene *h, long long *n)
{
    if (!h || !n) return VG_EINVAL;
    const SceneHost *S = scene0(h);
    if (!S) return fail(h, VG_EINVAL, "vg_num_interactions needs a one-scene handle");
    if (int rc = flush_rebuild(h, true)) return rc;
    if (h->n_iact < 0) {
        unsigned long long *d = nullptr, c[2] = {0, 0};
        if (dalloc(h, &d, 2)) return VG_ENOMEM;
        HIPCHK(hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), h->stream));
        hipLaunchKernelGGL(k_count_iact, dim3(nblk(S->n, 256)), dim3(256), 0, h->stream, (const unsigned *)h->ocount,
                           (const unsigned *)h->onbr, S->seg_begin, S->seg_end, h->ld, h->ko, d);
        HIPCHK(hipMemcpyAsync(c, d, sizeof(c), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipFree(d));
        h->n_iact = (long long)(c[0] / 2 + c[1]);   // sphere pairs appear in both spheres' lists
    }
    *n = h->n_iact;
    return 0;
}

extern "C" int vg_set_skin_auto(vg_scene *h, double floor, double cap)
{
    if (!h || !(floor >= 0.0) || !(cap >= floor)) return VG_EINVAL;
    h->skin_auto = true;
    h->skin_floor = floor;
    h->skin_cap = cap;
    return 0;
}

extern "C" int vg_set_skin(vg_scene *h, double skin)
{
    if (!h || !(skin >= 0.0)) return VG_EINVAL;
    h->skin_auto = false;
    HIPCHK(hipStreamSynchronize(h->stream));
    // every scene takes the skin (at least its bond reach) and relists at the next step
    for (int k = 0; k < h->scap; ++k) {
        SceneHost &S = h->sc[k];
        if (S.state == VG_SCENE_FREE || S.state == VG_SCENE_DONE) continue;
        S.skin = std::max(skin, S.bond_reach * 1.01);
        h->sdev[k].skin = S.skin;
        h->sdev[k].relist_half = 0.5 * S.skin * (1.0 - 1e-9);
    }
    if (upload_scenes(h)) return VG_EHIP;
    std::vector<int> r(3 * (size_t)h->scap);
    HIPCHK(hipMemcpy(r.data(), h->rflags, sizeof(int) * r.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < h->scap; ++k) r[3 * k + (h->step_count % 3)] = 1;
    HIPCHK(hipMemcpy(h->rflags, r.data(), sizeof(int) * r.size(), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int vg_set_precision(vg_scene *h, int precision)
{
    if (!h || (precision != VG_PRECISION_FP64 && precision != VG_PRECISION_MIXED)) return VG_EINVAL;
    h->precision = precision;
    return 0;
}

extern "C" int vg_get_precision(const vg_scene *h) { return h ? h->precision : -1; }

extern "C" double vg_get_skin(const vg_scene *h)
{
    if (!h) return 0.0;
    for (const SceneHost &S : h->sc)
        if (S.state == VG_SCENE_RUNNING || S.state == VG_SCENE_QUEUED || S.state == VG_SCENE_FINISHED) return S.skin;
    return 0.0;
}

static void status_of(const ActiveOut &A, vg_status &S)
{
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

static int get_active(vg_scene *h, int slot, long long *n, unsigned *lo, unsigned *hi, vg_status *st)
{
    HIPCHK(hipStreamSynchronize(h->stream));
    unsigned long long cnt = 0;
    std::vector<ActiveOut> ho;
    if (!h->batch && h->snap_valid) {
        cnt = (unsigned long long)h->snap_count;
        if (st) {
            ho.resize(std::min<size_t>(cnt, (size_t)std::max(*n, 0ll)));
            if (!ho.empty()) HIPCHK(hipMemcpy(ho.data(), h->snap, sizeof(ActiveOut) * ho.size(), hipMemcpyDeviceToHost));
        }
    } else {
        unsigned long long *dcnt = nullptr;
        unsigned *d_sid = nullptr, *d_wid = nullptr;
        if (dalloc(h, &dcnt, 1) || id_maps(h, slot, &d_sid, &d_wid)) return VG_ENOMEM;
        // after a plain step the entries hold the step's Backup as pending; at
        // a rebuild step (its rebuild deferred) the records are read before it,
        // as RebuildIactRecordTab reads them
        const bool backup = !h->rebuild_pending;
        if (export_active(h, slot, nullptr, 0, dcnt, d_sid, d_wid, backup)) return VG_EHIP;
        HIPCHK(hipMemcpyAsync(&cnt, dcnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (st) {
            const size_t cap = (size_t)std::max(*n, 0ll);
            ActiveOut *dout = nullptr;
            if (dalloc(h, &dout, cap)) return VG_ENOMEM;
            if (export_active(h, slot, dout, cap, dcnt, d_sid, d_wid, backup)) return VG_EHIP;
            ho.resize(std::min<size_t>(cnt, cap));
            if (!ho.empty())
                HIPCHK(hipMemcpyAsync(ho.data(), dout, sizeof(ActiveOut) * ho.size(), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            HIPCHK(hipFree(dout));
        }
        HIPCHK(hipFree(dcnt));
        HIPCHK(hipFree(d_sid));
        HIPCHK(hipFree(d_wid));
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
        lo[k] = ho[k].lo;
        hi[k] = ho[k].hi;
        status_of(ho[k], st[k]);
    }
    *n = (long long)ho.size();
    return 0;
}

extern "C" int vg_scene_num_do(vg_scene *h, int scene, int *n_do)
{
    if (!h || !n_do || scene < 0 || scene >= h->scap) return VG_EINVAL;
    const SceneHost &S = h->sc[scene];
    if (S.state == VG_SCENE_FREE) return fail(h, VG_EINVAL, "no scene in this slot");
    *n_do = S.n_do;
    return 0;
}

extern "C" int vg_scene_models(vg_scene *h, int scene, int *model_of)
{
    if (!h || !model_of || scene < 0 || scene >= h->scap) return VG_EINVAL;
    const SceneHost &S = h->sc[scene];
    if (S.state == VG_SCENE_FREE) return fail(h, VG_EINVAL, "no scene in this slot");
    std::copy(S.do_model.begin(), S.do_model.end(), model_of);
    return 0;
}

extern "C" int vg_get_active(vg_scene *h, long long *n, unsigned *lo, unsigned *hi, vg_status *st)
{
    if (!h || !n) return VG_EINVAL;
    if (!scene0(h)) return fail(h, VG_EINVAL, "vg_get_active needs a one-scene handle (vg_scene_get_active)");
    return get_active(h, 0, n, lo, hi, st);
}

extern "C" int vg_set_history(vg_scene *h, int n, const unsigned *lo, const unsigned *hi, const vg_status *st)
{
    if (!h || n < 0 || (n > 0 && (!lo || !hi || !st))) return VG_EINVAL;
    const SceneHost *S0 = scene0(h);
    if (!S0) return fail(h, VG_EINVAL, "vg_set_history applies to a one-scene handle");
    if (h->step_count != 0)
        return fail(h, VG_EINVAL, "histories are restored before the first step (SimMediator::Initiate)");
    if (n == 0) return 0;
    // records by key lo << 32 | hi (IactRecordTab is keyed by the DO-id pair)
    std::vector<std::pair<unsigned long long, int>> order((size_t)n);
    for (int k = 0; k < n; ++k) {
        if (lo[k] >= (unsigned)S0->n_do || hi[k] >= (unsigned)S0->n_do || lo[k] == hi[k])
            return fail(h, VG_EINVAL, "a record names an unknown DO or pairs a DO with itself");
        const unsigned a = std::min(lo[k], hi[k]), b = std::max(lo[k], hi[k]);
        order[(size_t)k] = {((unsigned long long)a << 32) | b, k};
    }
    std::stable_sort(order.begin(), order.end());
    std::vector<unsigned long long> keys((size_t)n);
    std::vector<RestoreRec> recs((size_t)n);
    for (int q = 0; q < n; ++q) {
        if (q > 0 && order[(size_t)q].first == order[(size_t)q - 1].first)
            return fail(h, VG_EINVAL, "two records for one pair");
        const vg_status &s = st[order[(size_t)q].second];
        RestoreRec &r = recs[(size_t)q];
        r = RestoreRec();
        r.flags = (s.contact ? F_CONTACT : 0u) | (s.last_contact ? F_LASTCONTACT : 0u) | (s.bond ? F_BOND : 0u) |
                  (s.last_bond ? F_LASTBOND : 0u);
        r.kn = s.kn;
        r.init_vel = s.init_vel;
        for (int c = 0; c < 3; ++c) r.shear[c] = s.shear[c];
        keys[(size_t)q] = order[(size_t)q].first;
    }
    unsigned long long *d_keys = nullptr;
    RestoreRec *d_recs = nullptr;
    unsigned *d_sid = nullptr, *d_wid = nullptr;
    int rc = 0;
    if (dalloc(h, &d_keys, (size_t)n) || dalloc(h, &d_recs, (size_t)n) || id_maps(h, 0, &d_sid, &d_wid)) rc = VG_ENOMEM;
    if (!rc && (hipMemcpy(d_keys, keys.data(), sizeof(unsigned long long) * n, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d_recs, recs.data(), sizeof(RestoreRec) * n, hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(h, VG_EHIP, "record upload failed");
    if (!rc && S0->n > 0) {
        hipLaunchKernelGGL(k_restore_history, dim3(nblk(S0->n, 128)), dim3(128), 0, h->stream, h->geo[h->cur],
                           h->ell[h->ecur], S0->seg_begin, S0->seg_end, h->ld, h->ki, (const unsigned *)d_sid,
                           (const unsigned *)d_wid, (const unsigned long long *)d_keys, (const RestoreRec *)d_recs,
                           (long long)n, (const ModelDev *)h->models, (const SceneDev *)h->scenes,
                           (const LawDev *)h->laws, (const WallDev *)h->walls);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess)
            rc = fail(h, VG_EHIP, "history restore kernel failed");
    }
    for (void *p : {(void *)d_keys, (void *)d_recs, (void *)d_sid, (void *)d_wid})
        if (p) (void)hipFree(p);
    h->snap_valid = false;
    h->ent_valid = false;
    return rc;
}

// ---------------------------------------------------------------------------
// Metrics, timing, counters
// ---------------------------------------------------------------------------

static int system_energy(vg_scene *h, int slot, double *out)
{
    const SceneHost &S = h->sc[slot];
    const int n = S.n;
    const int nb = (int)std::max(1u, std::min(1024u, nblk(n, 256)));
    double *avg = h->en_out + 16;
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_energy_part, dim3(nb), dim3(256), 0, h->stream, h->geo[h->cur], S.seg_begin, S.seg_end,
                           (const double2 *)h->mm, S.g[0], S.g[1], S.g[2], (const double *)avg, pass, h->en_part);
        hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, h->stream, (const double *)h->en_part, nb, n, pass,
                           avg, h->en_out);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, h->en_out, 16 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int vg_system_energy(vg_scene *h, double *out)
{
    if (!h || !out) return VG_EINVAL;
    if (!scene0(h)) return fail(h, VG_EINVAL, "vg_system_energy needs a one-scene handle (vg_scene_energy)");
    return system_energy(h, 0, out);
}

// ---------------------------------------------------------------------------
// Knight analyses (src/vedo/knight) of a scene's current state
// ---------------------------------------------------------------------------

static int check_slot(vg_scene *h, int scene);

static int knight_scene(vg_scene *h, int scene, const SceneHost **out)
{
    if (int rc = check_slot(h, scene)) return rc;
    const SceneHost &S = h->sc[scene];
    if (S.state != VG_SCENE_RUNNING && S.state != VG_SCENE_FINISHED)
        return fail(h, VG_EINVAL, "scene is not on the device");
    *out = &S;
    return 0;
}

static KnightBox knight_box(const double *b)
{
    KnightBox B;
    B.lo[0] = b[0]; B.hi[0] = b[1];
    B.lo[1] = b[2]; B.hi[1] = b[3];
    B.lo[2] = b[4]; B.hi[2] = b[5];
    return B;
}

extern "C" int vg_scene_coordination_number(vg_scene *h, int scene, const double *box, double *avg)
{
    const SceneHost *S = nullptr;
    if (!box || !avg) return VG_EINVAL;
    if (int rc = knight_scene(h, scene, &S)) return rc;
    if (!S->all_sphere_laws)
        return fail(h, VG_EUNSUPPORTED, "coordination number needs an IactModel for every pair of sphere groups");
    if (int rc = flush_rebuild(h, true)) return rc;   // the touching pairs of the rebuilt candidate list
    unsigned long long *acc = nullptr, c[2] = {0, 0};
    if (dalloc(h, &acc, 2)) return VG_ENOMEM;
    int rc = 0;
    if (hipMemsetAsync(acc, 0, sizeof(c), h->stream) != hipSuccess) rc = fail(h, VG_EHIP, "memset failed");
    if (!rc && S->n > 0)
        hipLaunchKernelGGL(k_coordination, dim3(nblk(S->n, 256)), dim3(256), 0, h->stream, h->geo[h->cur], S->seg_begin,
                           S->seg_end, h->ld, (const unsigned *)h->ocount, (const unsigned *)h->onbr, knight_box(box),
                           acc);
    if (!rc && (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(c, acc, sizeof(c), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
                hipStreamSynchronize(h->stream) != hipSuccess))
        rc = fail(h, VG_EHIP, "coordination kernel failed");
    (void)hipFree(acc);
    if (rc) return rc;
    *avg = (double)(unsigned)c[0] / (double)c[1];   // total / number of the box's spheres (NaN for none)
    return 0;
}

// njr::RandomGenerator (RandomGenerator.cpp:70-126), the knight tool's
// generator: a subtractive lagged-Fibonacci sequence modulo 1e9 over 55
// words, seeded by sprand(seed) and 165 discarded draws.
namespace {
struct NjrRng {
    int a = 0, b = 24, arr[55];
    explicit NjrRng(int seed)
    {
        int last = seed, next = 1;
        arr[0] = seed;
        for (int i = 1; i < 55; i++) {
            const int ii = (21 * i) % 55;
            arr[ii] = next;
            next = last - next;
            if (next < 0) next += 1000000000;
            last = arr[ii];
        }
        a = 0;
        b = 24;
        for (int i = 0; i < 165; i++) draw();
    }
    int draw()
    {
        if ((a--) == 0) a = 54;
        if ((b--) == 0) b = 54;
        int t = arr[a] - arr[b];
        if (t < 0) t += 1000000000;
        arr[a] = t;
        return t;
    }
    double uniform(double lo, double hi)   // GeneratingDouble
    {
        if (lo > hi) std::swap(lo, hi);
        else if (lo == hi) return hi;
        return (double)draw() / 1000000000 * (hi - lo) + lo;
    }
};
}  // namespace

extern "C" int vg_scene_packing_density(vg_scene *h, int scene, const double *box, int n_points, int seed,
                                        double *percent)
{
    const SceneHost *S = nullptr;
    if (!box || !percent || n_points <= 0) return VG_EINVAL;
    if (int rc = knight_scene(h, scene, &S)) return rc;
    std::vector<double> pts((size_t)3 * n_points);
    NjrRng rng(seed);
    for (int k = 0; k < n_points; ++k) {
        pts[3 * (size_t)k] = rng.uniform(box[0], box[1]);
        pts[3 * (size_t)k + 1] = rng.uniform(box[2], box[3]);
        pts[3 * (size_t)k + 2] = rng.uniform(box[4], box[5]);
    }
    double *dp = nullptr;
    unsigned long long *hits = nullptr, hh = 0;
    if (dalloc(h, &dp, pts.size()) || dalloc(h, &hits, 1)) {
        if (dp) (void)hipFree(dp);
        return VG_ENOMEM;
    }
    int rc = 0;
    const SceneDev &D = h->sdev[scene];
    if (hipMemcpyAsync(dp, pts.data(), sizeof(double) * pts.size(), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemsetAsync(hits, 0, sizeof(hh), h->stream) != hipSuccess)
        rc = fail(h, VG_EHIP, "upload failed");
    if (!rc && S->n > 0)
        hipLaunchKernelGGL(k_density_hits, dim3(nblk(n_points, 256)), dim3(256), 0, h->stream, h->geo[h->cur], D.G,
                           (long long)D.cell_off, (const int *)h->cstart, (const int *)h->cend, (const double *)dp,
                           n_points, knight_box(box), hits);
    if (!rc && (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(&hh, hits, sizeof(hh), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
                hipStreamSynchronize(h->stream) != hipSuccess))
        rc = fail(h, VG_EHIP, "density kernel failed");
    (void)hipFree(dp);
    (void)hipFree(hits);
    if (rc) return rc;
    // hit / (float)RandomHitTestNum * 100.0, formed in float as the reference does
    const float f = (float)(unsigned)hh / (float)n_points;
    *percent = (double)f * 100.0;
    return 0;
}

extern "C" int vg_scene_packing_density_noalloc(vg_scene *h, int scene, const double *box, int n_points, int seed,
                                                double *percent)
{
    const SceneHost *S = nullptr;
    if (!box || !percent || n_points <= 0) return VG_EINVAL;
    if (int rc = knight_scene(h, scene, &S)) return rc;
    if (n_points > SCRATCH_POINTS) {
        char msg[128];
        std::snprintf(msg, sizeof(msg), "vg_scene_packing_density_noalloc takes at most %d points (%d asked)",
                      SCRATCH_POINTS, n_points);
        return fail(h, VG_EINVAL, msg);
    }
    double *pts = scratch_of(h->device);
    if (!pts) return fail(h, VG_EHIP, "scratch array not set up on this device");
    unsigned long long *hits = reinterpret_cast<unsigned long long *>(pts + 3 * (size_t)SCRATCH_POINTS);
    // the counter is cleared by a kernel (no memset node if the caller captures the stream)
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, h->stream, reinterpret_cast<int *>(hits), 2ll, 0);
    // the points, drawn on the host in the reference's order, go up through a
    // stack buffer chunk by chunk (each copy completes before the buffer is refilled)
    constexpr int CHUNK = 2048;
    double buf[3 * CHUNK];
    NjrRng rng(seed);
    for (int k0 = 0; k0 < n_points; k0 += CHUNK) {
        const int m = std::min(CHUNK, n_points - k0);
        for (int k = 0; k < m; ++k) {
            buf[3 * k] = rng.uniform(box[0], box[1]);
            buf[3 * k + 1] = rng.uniform(box[2], box[3]);
            buf[3 * k + 2] = rng.uniform(box[4], box[5]);
        }
        if (hipMemcpyAsync(pts + 3 * (size_t)k0, buf, sizeof(double) * 3 * m, hipMemcpyHostToDevice, h->stream) !=
                hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess)
            return fail(h, VG_EHIP, "upload failed");
    }
    const SceneDev &D = h->sdev[scene];
    if (S->n > 0)
        hipLaunchKernelGGL(k_density_hits, dim3(nblk(n_points, 256)), dim3(256), 0, h->stream, h->geo[h->cur], D.G,
                           (long long)D.cell_off, (const int *)h->cstart, (const int *)h->cend, (const double *)pts,
                           n_points, knight_box(box), hits);
    unsigned long long hh = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&hh, hits, sizeof(hh), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return fail(h, VG_EHIP, "density kernel failed");
    // hit / (float)RandomHitTestNum * 100.0, formed in float as the reference does
    const float f = (float)(unsigned)hh / (float)n_points;
    *percent = (double)f * 100.0;
    return 0;
}

extern "C" int vg_scene_granular_temperature(vg_scene *h, int scene, int model, const double *center,
                                             const double *ox, const double *oz, const double *size, double *out)
{
    const SceneHost *S = nullptr;
    if (!center || !ox || !oz || !size || !out) return VG_EINVAL;
    if (int rc = knight_scene(h, scene, &S)) return rc;
    if (model < 0 || model >= S->n_models) return fail(h, VG_EINVAL, "unknown DOModel");
    if (!S->model_is_sphere[(size_t)model])
        return fail(h, VG_EUNSUPPORTED, "granular temperature of sphere models only");
    GtRegion q;
    q.model = S->model_off + model;
    q.c = mk(center[0], center[1], center[2]);
    q.ox = mk(ox[0], ox[1], ox[2]);
    q.oz = mk(oz[0], oz[1], oz[2]);
    q.oy = cross(q.oz, q.ox);
    q.w = size[0]; q.l = size[1]; q.ht = size[2];
    const int nb = (int)std::max(1u, std::min(512u, nblk(S->n, 256)));
    double *part = nullptr, *davg = nullptr;
    if (dalloc(h, &part, (size_t)7 * nb) || dalloc(h, &davg, 6)) {
        if (part) (void)hipFree(part);
        return VG_ENOMEM;
    }
    std::vector<double> hp((size_t)7 * nb);
    double sums[7] = {0, 0, 0, 0, 0, 0, 0}, avg[6] = {0, 0, 0, 0, 0, 0}, dev[2] = {0, 0};
    int rc = 0;
    for (int pass = 0; pass < 2 && !rc; ++pass) {
        if (pass == 1 && hipMemcpyAsync(davg, avg, sizeof(avg), hipMemcpyHostToDevice, h->stream) != hipSuccess) {
            rc = fail(h, VG_EHIP, "upload failed");
            break;
        }
        hipLaunchKernelGGL(k_gt_part, dim3(nb), dim3(256), 0, h->stream, h->geo[h->cur], S->seg_begin, S->seg_end, q,
                           pass, (const double *)davg, part);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            rc = fail(h, VG_EHIP, "granular temperature kernel failed");
            break;
        }
        for (int blk = 0; blk < nb; ++blk)   // block sums in block order
            for (int k = 0; k < 7; ++k) {
                if (pass == 0) sums[k] += hp[(size_t)7 * blk + k];
                else if (k < 2) dev[k] += hp[(size_t)7 * blk + k];
            }
        if (pass == 0 && sums[6] != 0.0) {
            const double inv = 1.0 / sums[6];   // 1.0 / Counter * sum (Knight.cpp:1186-1187)
            for (int k = 0; k < 6; ++k) avg[k] = inv * sums[k];
        }
    }
    (void)hipFree(part);
    (void)hipFree(davg);
    if (rc) return rc;
    const double n = sums[6];
    double tv = dev[0], tw = dev[1];
    if (n != 0.0) {
        tv = tv / n / 3.0;
        tw = tw / n / 3.0;
    }
    out[0] = n;
    out[1] = std::sqrt((avg[0] * avg[0] + avg[1] * avg[1]) + avg[2] * avg[2]);
    out[2] = tv;
    out[3] = std::sqrt((avg[3] * avg[3] + avg[4] * avg[4]) + avg[5] * avg[5]);
    out[4] = tw;
    return 0;
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
    for (auto e : h->rev) HIPCHK(hipEventDestroy(e));
    h->kev.clear();
    h->rev.clear();
    h->kev_used = 0;
    h->rev_used = 0;
    for (int k = 0; k < 2 * max_launches; ++k) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        h->kev.push_back(e);
    }
    // a rebuild comes at most every step
    const int max_rebuilds = std::min(max_launches, 4096);
    for (int k = 0; k < 2 * max_rebuilds; ++k) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        h->rev.push_back(e);
    }
    return 0;
}

extern "C" int vg_rebuild_timing(vg_scene *h, double *ms, int *rebuilds)
{
    if (!h) return VG_EINVAL;
    HIPCHK(hipStreamSynchronize(h->stream));
    double tot = 0.0;
    for (int k = 0; k < h->rev_used; ++k) {
        float m = 0.0f;
        HIPCHK(hipEventElapsedTime(&m, h->rev[2 * k], h->rev[2 * k + 1]));
        tot += m;
    }
    if (ms) *ms = tot;
    if (rebuilds) *rebuilds = h->rev_used;
    return 0;
}

extern "C" int vg_check_lists(vg_scene *h, long long *counts)
{
    if (!h || !counts) return VG_EINVAL;
    HIPCHK(hipStreamSynchronize(h->stream));
    unsigned long long *d = nullptr, c[11] = {0};
    if (dalloc(h, &d, 11)) return VG_ENOMEM;
    int rc = 0;
    if (hipMemsetAsync(d, 0, sizeof(c), h->stream) != hipSuccess) rc = fail(h, VG_EHIP, "memset failed");
    if (!rc && h->n > 0)
        hipLaunchKernelGGL(k_check_lists, dim3(nblk(h->n, 256)), dim3(256), 0, h->stream, h->geo[h->cur],
                           h->ell[h->ecur], h->n, h->ld, h->ki, h->ko, (const unsigned *)h->ocount,
                           (const unsigned *)h->onbr, (const ModelDev *)h->models, h->models_cap,
                           (const SceneDev *)h->scenes, h->scap, (const WallDev *)h->walls, h->wcap,
                           h->any_constrained ? (const int *)h->cw_w : (const int *)nullptr,
                           (const int *)h->cw_p, h->cw_cap, d);
    if (!rc && (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(c, d, sizeof(c), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
                hipStreamSynchronize(h->stream) != hipSuccess))
        rc = fail(h, VG_EHIP, "list check failed");
    (void)hipFree(d);
    if (rc) return rc;
    for (int k = 0; k < 11; ++k) counts[k] = (long long)c[k];
    return 0;
}

extern "C" int vg_stats(vg_scene *h, long long *rebuilds, long long *relists, int *max_fill, int *capacity,
                        long long *contact_steps)
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
    if (contact_steps) {
        std::vector<unsigned long long> c(nblk(h->ld, 64));
        if (!c.empty())
            HIPCHK(hipMemcpy(c.data(), h->contacts, sizeof(unsigned long long) * c.size(), hipMemcpyDeviceToHost));
        unsigned long long t = 0;
        for (unsigned long long x : c) t += x;
        *contact_steps = (long long)t;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Batch: per-scene queries
// ---------------------------------------------------------------------------

static int check_slot(vg_scene *h, int scene)
{
    if (!h) return VG_EINVAL;
    if (scene < 0 || scene >= h->scap || h->sc[scene].state == VG_SCENE_FREE)
        return fail(h, VG_EINVAL, "unknown scene");
    return 0;
}

extern "C" int vg_scene_status(vg_scene *h, int scene, int *state, long long *steps_done, double *t)
{
    if (int rc = check_slot(h, scene)) return rc;
    const SceneHost &S = h->sc[scene];
    if (state) *state = S.state == VG_SCENE_RETIRING ? VG_SCENE_FINISHED : S.state;
    if (steps_done) *steps_done = S.steps_done;
    if (t) *t = S.t;
    return 0;
}

extern "C" int vg_scene_get_state(vg_scene *h, int scene, double *pos, double *vel, double *ox, double *oz,
                                  double *av)
{
    if (int rc = check_slot(h, scene)) return rc;
    SceneHost &S = h->sc[scene];
    if (S.state == VG_SCENE_QUEUED) return fail(h, VG_EINVAL, "scene not placed yet");
    if (S.state == VG_SCENE_DONE) {
        split_state(S.result, S.n_do, pos, vel, ox, oz, av, nullptr, nullptr);
        return 0;
    }
    std::vector<double> hb;
    if (fetch_scene_state(h, scene, hb)) return VG_EHIP;
    split_state(hb, S.n_do, pos, vel, ox, oz, av, nullptr, nullptr);
    return 0;
}

extern "C" int vg_scene_get_active(vg_scene *h, int scene, long long *n, unsigned *lo, unsigned *hi, vg_status *st)
{
    if (int rc = check_slot(h, scene)) return rc;
    if (!n) return VG_EINVAL;
    const SceneHost &S = h->sc[scene];
    if (S.state != VG_SCENE_RUNNING && S.state != VG_SCENE_FINISHED)
        return fail(h, VG_EINVAL, "scene is not on the device");
    return get_active(h, scene, n, lo, hi, st);
}

extern "C" int vg_scene_energy(vg_scene *h, int scene, double *out)
{
    if (int rc = check_slot(h, scene)) return rc;
    if (!out) return VG_EINVAL;
    const SceneHost &S = h->sc[scene];
    if (S.state != VG_SCENE_RUNNING && S.state != VG_SCENE_FINISHED)
        return fail(h, VG_EINVAL, "scene is not on the device");
    return system_energy(h, scene, out);
}

extern "C" int vg_scene_set_wall_rotation(vg_scene *h, int scene, int n, const unsigned *ids, const double *c,
                                          const double *a, double omega, double t_ref)
{
    if (int rc = check_slot(h, scene)) return rc;
    SceneHost &S = h->sc[scene];
    if (!ids || !c || !a || n < 0) return VG_EINVAL;
    if (S.state == VG_SCENE_QUEUED) {
        // not placed yet: the prescription is applied when the scene is placed
        const double al = olen(a);
        if (S.kin0.empty()) {
            S.kin0.assign(S.nw, WallKin());
            S.kin0_set.assign(S.nw, 0);
        }
        for (int k = 0; k < n; ++k) {
            if (ids[k] >= (unsigned)S.n_do || S.do_to_wall[ids[k]] < 0)
                return fail(h, VG_EINVAL, "wall rotation applies to walls only");
            const int w = S.do_to_wall[ids[k]] - S.w_begin;
            WallKin &K = S.kin0[w];
            K.c = mk(c[0], c[1], c[2]);
            K.a = mk(a[0] / al, a[1] / al, a[2] / al);
            K.omega = omega;
            K.t_ref = t_ref;
            S.kin0_set[w] = 1;
        }
        return 0;
    }
    if (S.state != VG_SCENE_RUNNING && S.state != VG_SCENE_FINISHED)
        return fail(h, VG_EINVAL, "scene is not on the device");
    return set_rotation(h, scene, n, ids, c, a, omega, t_ref);
}

extern "C" int vg_retire(vg_scene *h, int scene)
{
    if (int rc = check_slot(h, scene)) return rc;
    SceneHost &S = h->sc[scene];
    if (S.state == VG_SCENE_DONE || S.state == VG_SCENE_FAILED) {
        S = SceneHost();   // forget the kept result
        return 0;
    }
    if (S.state == VG_SCENE_QUEUED) {
        for (int w = 0; w < S.nw; ++w) h->wall_used[S.w_begin + w] = 0;
        S = SceneHost();
        return 0;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    // leaves the device at the next boundary, without a kept result
    S.result.clear();
    S.state = VG_SCENE_RETIRING;
    h->sdev[scene].active = 0;
    return upload_scenes(h);
}

extern "C" int vg_batch_info(vg_scene *h, long long *spheres, int *scenes_on_device, int *scenes_queued,
                             unsigned long long *steps)
{
    if (!h) return VG_EINVAL;
    int on = 0, q = 0;
    for (const SceneHost &S : h->sc) {
        if (S.state == VG_SCENE_RUNNING || S.state == VG_SCENE_FINISHED || S.state == VG_SCENE_RETIRING) ++on;
        if (S.state == VG_SCENE_QUEUED) ++q;
    }
    if (spheres) *spheres = h->n;
    if (scenes_on_device) *scenes_on_device = on;
    if (scenes_queued) *scenes_queued = q;
    if (steps) *steps = h->step_count;
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
