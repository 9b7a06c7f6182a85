// vg_kernels.h -- kernel argument blocks shared by the kernels and the host.
#pragma once

#include "vg_device.h"

namespace vg {

struct Geo {   // one buffer set of sphere state (see vg_device.h)
    double4 *pos, *vel, *avl;
    double *ori;     // [6][LD]
    unsigned *oid;   // sphere rank within its scene: position among the scene's spheres in DO-id order
    int *mdl;
};

struct Ell {
    unsigned char *hdr;
    unsigned *ent;    // [2][KI][LD]
    unsigned char *slot;   // [2][KI][LD]: the entry's history slot while its Contact flag is set
    double4 *hd;      // [KI][LD] per history slot: shear force xyz, then Kn (ISwH) / initial velocity
                      // (ISwHertz), zero for the spring-dashpot laws; two 16-byte accesses
};

struct Tables {
    const ModelDev *models;
    const LawDev *laws;   // per scene [G][G] at SceneDev::law_off
};

// Cell grid of a rebuild (LeapConsultant.h:103-144).
struct GridDev {
    double xmin, ymin, zmin, zr;
    double safe_v, safe_g;   // dt_up and 0.5*dt_up*dt_up*|g|
    double sf;
    int nx, ny, nz;          // dense dims (reference ncel + 3)
    int perm[3];             // comparator axis priority
    // The 14 cell groups of a sphere's traversal in comparator order, 4 bits
    // each from bit 0: 13 is the sphere's own cell, k < 13 the cell at
    // -NB13[k] (whose half shell holds the own cell).  The offsets alone
    // decide their comparator order, so it is the same for every cell.
    unsigned long long order;
};

// The 13-cell half shell of the Leap traversal (LeapConsultant.h:207-444).
#define VG_NB13_INIT                                                                                             \
    {                                                                                                           \
        {-1, -1, -1}, {-1, -1, 0}, {-1, -1, 1}, {0, -1, -1}, {0, -1, 0}, {0, -1, 1}, {1, -1, -1}, {1, -1, 0},  \
            {1, -1, 1}, {-1, 0, -1}, {-1, 0, 0}, {0, 0, -1}, {1, 0, -1}                                         \
    }

// One scene of a batch: its SystemParameter scalars, IactModel table, walls,
// grid of the last rebuild and state.  A scene's spheres occupy one
// contiguous run of slots, in scene-slot order.
struct SceneDev {
    double dt, gx, gy, gz;
    double skin, relist_half;
    GridDev G;
    long long cell_off;      // first cell of the scene's grid in cstart / cend
    int law_off, n_groups;   // laws[law_off + gi * n_groups + gj]
    int w_begin, w_end;      // the scene's wall slots
    int rank_off;            // offset of the scene's ranks in the sort of the next rebuild
    int sid_off;             // offset of the scene's sphere rank -> DO id map
    int active;              // 1 stepping, 0 frozen (state held), -1 retired (dropped at the rebuild)
    int zon;                 // zone of interest: bit d set where axis d is bounded (0: none)
    double zlo[3], zhi[3];   // its bounds (SystemParameter ZoneOfInterest, Boundary.cpp:79-191)
};

// A sphere erased by the zone of interest at a rebuild carries, in place of
// its rank, OID_DEAD | its rank in the dropped region of the rebuild's sort.
constexpr unsigned OID_DEAD = 0xC0000000u;

struct StepArgs {
    int n, ld, nw, ki;       // live sphere slots, leading dimension, wall slots, list capacity
    Geo in, out;
    Ell ell;
    const unsigned *o_cnt;   // candidate-list length per sphere [LD]
    const unsigned *o_nbr;   // candidate list, ELL [KO][LD]
    int ko;
    Tables t;
    const SceneDev *scenes;
    int n_scenes;
    const WallDev *walls;
    int *rflags;             // relist ring per scene [S][3]
    int *flags;              // [3] error bits, [4] max fill, [5] candidate overflow
    unsigned long long *relists;
    unsigned long long *contacts;   // per-wavefront running count of solved active contact pairs
    int step;
    int single_group;        // every scene's spheres share one group: no partner group lookup
    int no_bond;             // no bonded law: surely separated pairs skip the contact geometry
    int group0, law_ss;      // one scene whose spheres share one group: that group and its sphere-sphere law
    int any_fext;            // a sphere model has an external force
    // constrained walls: a sphere's constrained walls and their pair slots [CW_SLOTS][LD],
    // each pair's master impulse and the step it was solved in
    const int *cw_w, *cw_p;
    double *cw_val;
    unsigned *cw_stamp;
    unsigned stamp;
    // record path
    double *rec_imp;         // [6][LD] last impulse / angular impulse per sphere
    double *rec;             // [KI][13][LD] point, dir, to_master, ang_to_master, overlap
    const double4 *holes;    // per wall slot: hole radius, x / y offset, 1 for a holed plate (LS_EXT kernels)
    const PbcDev *pbc;       // periodic boundary of the (single) scene, or null (LS_EXT kernels)
};

enum { ERR_INNER_CAP = 1, ERR_OUTER_CAP = 2, ERR_CW = 4 };

// Neighbour-list entries per sphere: the list header keeps the count in 7
// bits.  Lists longer than 64 entries take the step kernel's wide variant
// (two 64-bit words for its entry and history-slot sets).
constexpr int KI_MAX = 127;

// Constrained walls a sphere may have among its candidates (pair slots per sphere).
constexpr int CW_SLOTS = 4;

// Step kernel launch shape: threads per workgroup (spheres per tile) and the
// workgroups per CU that its registers and LDS allow.
#ifndef VG_STEP_NT
#define VG_STEP_NT 256
#endif
#ifndef VG_STEP_WG
#define VG_STEP_WG (768 / VG_STEP_NT)
#endif
constexpr int STEP_NT = VG_STEP_NT, STEP_WG_PER_CU = VG_STEP_WG;
static_assert(STEP_NT % 64 == 0 && STEP_NT <= 256, "the step kernel's block is 1 to 4 wavefronts");

// Kinematic prescription of a wall: rigid rotation about axis (c, a).
struct WallKin {
    d3 c, a;
    double omega, t_ref;
};

struct OuterArgs {
    Geo g;
    int n, ld;
    const SceneDev *scenes;
    const int *cstart, *cend;
    const double *safe;
    const WallDev *walls;
    Tables t;
    int single_group;        // every scene's spheres share one group
    unsigned *count;         // candidates per sphere [LD]
    unsigned *nbr;           // ELL [KO][LD]
    int ko;
    int ki;                  // neighbour-list capacity: spheres with more entries within the skin raise flags[6]
    int *flags;              // [5]: largest candidate count seen; [6]: largest neighbour count above ki
    int *ovf;                // per scene: a sphere has more candidates than ko (batches; may be null)
    const double4 *holes;    // per wall slot: hole radius, x / y offset, 1 for a holed plate (HOLE kernels)
};

// candidate list under a periodic boundary (a handle of one scene)
struct PbcArgs {
    Geo g;
    int n, ld;
    const SceneDev *scene;
    const int *cstart, *cend;
    const double *safe;
    const WallDev *walls;
    const double4 *holes;
    const PbcDev *pbc;
    Tables t;
    const unsigned long long *ckey;   // occupied cells' comparator keys, sorted
    const unsigned *cfirst;           // first sphere of each sorted cell
    int ncell;
    unsigned *pcount;                 // pairs per cell
    const unsigned *poff;             // first pair of each cell
    uint2 *pairs;                     // (sphere, partner | marks) in reference order
    int nbs;                          // NBSConsultant's traversal instead of Leap's
    unsigned nc[3];                   // NBS: zones per axis (the traversal's ranges)
};

struct InnerArgs {
    Geo g, gold;             // new order / previous order
    Ell ell, eold;
    int n, ld, ki, ki_old;   // capacities of the new and the previous lists
    const unsigned *perm;    // new index -> previous index
    const unsigned *cnt, *nbr;
    int ko;
    const SceneDev *scenes;
    const ModelDev *models;
    const WallDev *walls;
    int *flags;
};

struct ActiveOut {
    unsigned lo, hi, flags, pad;
    double hist[4];
    double rec[13];
};

template <bool REC, int LS, bool MIX, bool UNI, int W = 1> __global__ void k_step(StepArgs a);
__global__ void k_walls(WallDev *walls, int nw, const SceneDev *scenes, double *stime, const double *assigned,
                        const WallDev *ref, const WallKin *kin, const double *cw_imp, int *rflags, int step);
__global__ void k_cw_count(int n, int ld, int ko, const unsigned *cnt, const unsigned *nbr, const WallDev *walls,
                           unsigned *out, int *flags);
__global__ void k_cw_emit(int n, int ld, int ko, const unsigned *cnt, const unsigned *nbr, const WallDev *walls,
                          Geo g, const ModelDev *models, const SceneDev *scenes, const unsigned *off,
                          unsigned long long *key, unsigned *val);
__global__ void k_cw_finalize(long long np, const unsigned long long *key, const unsigned *val, int ld, int *cw_w,
                              int *cw_p, unsigned *stamp, int *beg, int *end);
__global__ void k_cw_sum(const WallDev *walls, int nw, const SceneDev *scenes, const int *beg, const int *end,
                         const double *val, const unsigned *stamp, unsigned cur, double *imp);
__global__ void k_wall_path_reset(WallDev *walls, int nw);
__global__ void k_ring_clear(int *rflags, int scap, int r);
__global__ void k_bounds(Geo g, int n, const SceneDev *scenes, const ModelDev *models, unsigned long long *red,
                         unsigned *inv, double *xs, double *ys, double *zs, unsigned *sid, unsigned retired);
struct QselTarget {
    unsigned long long prefix;   // key bits settled so far
    unsigned long long rank;     // rank still to find among the keys with that prefix
};
__global__ void k_qsel_hist(Geo g, int n, const unsigned *sid, const int *qmap, int n_slots, const QselTarget *tg,
                            int pass, unsigned *hist);
__global__ void k_qsel_pick(QselTarget *tg, int m, int pass, unsigned *hist, double *out);
__global__ void k_quantiles(const double *xs, const double *ys, const double *zs, const long long *idx, int m,
                            double *out);
__global__ void k_cellkey(Geo g, int n_ranks, const SceneDev *scenes, const ModelDev *models, const unsigned *inv,
                          unsigned long long retired_key, int hbits, unsigned long long *keys, unsigned *vals);
__global__ void k_reorder(Geo src, Geo dst, int n, int ld, const unsigned *perm, const SceneDev *scenes,
                          const ModelDev *models, double *safe, int *lin, const double *imp_src, double *imp_dst);
__global__ void k_fill_int(int *p, long long n, int v);
__global__ void k_cellstart(const int *lin, int n, int *cstart, int *cend);
template <bool HOLE> __global__ void k_outer(OuterArgs a);
__global__ void k_pbc_cells(Geo g, int n, const int *lin, const SceneDev *scene, unsigned long long *ckey,
                            unsigned *cfirst, unsigned *ncell);
__global__ void k_pbc_count(PbcArgs a);
__global__ void k_pbc_emit(PbcArgs a);
__global__ void k_pbc_entries(const uint2 *pairs, unsigned np, int ld, unsigned *key, unsigned *val);
__global__ void k_pbc_segstart(const unsigned *key, unsigned ne, unsigned *first);
__global__ void k_pbc_lists(const unsigned *key, const unsigned *val, unsigned ne, const uint2 *pairs,
                            const unsigned *first, int ld, int ko, int ki, unsigned *count, unsigned *nbr,
                            int *flags);
__global__ void k_count_iact(const unsigned *cnt, const unsigned *nbr, int b, int e, int ld, int ko,
                             unsigned long long *out);
__global__ void k_inner_build(InnerArgs a);
__global__ void k_hdr_zero(unsigned char *hdr, int b, int e);
struct RestoreRec;
__global__ void k_zoi_mark(Geo g, int n, const SceneDev *scenes, const ModelDev *models, unsigned char *dead,
                           int *count);
__global__ void k_zoi_rank(Geo g, int b, int e, const unsigned char *dead, unsigned *dead_by_rank);
__global__ void k_zoi_newoid(Geo g, int b, int e, const unsigned char *dead, const unsigned *pre, unsigned dead_base);
__global__ void k_zoi_records(const ActiveOut *in, long long n, const unsigned *jump, unsigned long long *okey,
                              unsigned long long *nkey, unsigned *prio_src, RestoreRec *recs);
__global__ void k_iota(unsigned *p, long long n);
__global__ void k_gather_u64(const unsigned long long *src, const unsigned *idx, long long n, unsigned long long *dst);
__global__ void k_gather_restore(const RestoreRec *src, const unsigned *idx, long long n, RestoreRec *dst);
__global__ void k_zoi_prio(const unsigned long long *okey_sorted, const unsigned *idx_sorted,
                           const unsigned long long *nkey, long long n, unsigned *prio);
__global__ void k_restore_history(Geo g, Ell ell, int b, int e, int ld, int ki, const unsigned *sphere_id,
                                  const unsigned *wall_id, const unsigned long long *keys, const RestoreRec *recs,
                                  long long nrec, const ModelDev *models, const SceneDev *scenes, const LawDev *laws,
                                  const WallDev *walls);
__global__ void k_scatter_state(Geo g, int b, int e, int ld, const double *rec_imp, double *out);
template <bool EXT>
__global__ void k_export_active(Geo g, Ell ell, int b, int e, int ld, int ki, const unsigned *o_cnt,
                                const unsigned *o_nbr, const WallDev *walls, const unsigned *wall_id,
                                const unsigned *sphere_id, const double *rec, ActiveOut *out, unsigned long long cap,
                                unsigned long long *counter, int backup, int no_bond, const double4 *holes,
                                const PbcDev *pbc);
__global__ void k_check_lists(Geo g, Ell ell, int n, int ld, int ki, int ko, const unsigned *o_cnt,
                              const unsigned *o_nbr, const ModelDev *models, int n_models, const SceneDev *scenes,
                              int scap, const WallDev *walls, int wcap, const int *cw_w, const int *cw_p,
                              long long cw_cap, unsigned long long *out);
__global__ void k_wall_impulse(Ell ell, int b, int e, int ld, int ki, int w0, const double *rec, int no_bond, double *out);
__global__ void k_assign_spheres(Geo g, int b, int e, int ld, const double *buf, double *rec_imp);
// Knight analyses (src/vedo/knight) of a scene's current state.
struct KnightBox {
    double lo[3], hi[3];   // xmin, ymin, zmin / xmax, ymax, zmax
};
__global__ void k_coordination(Geo g, int b, int e, int ld, const unsigned *o_cnt, const unsigned *o_nbr, KnightBox box,
                               unsigned long long *acc);
__global__ void k_density_hits(Geo g, GridDev G, long long cell_off, const int *cstart, const int *cend,
                               const double *pts, int np, KnightBox box, unsigned long long *hits);
struct GtRegion {
    int model;                 // batch-wide DOModel index
    d3 c, ox, oy, oz;          // GSRectangle centre and axes (oy = oz x ox)
    double w, l, ht;           // width, length, height
};
__global__ void k_gt_part(Geo g, int b, int e, GtRegion q, int pass, const double *avg, double *part);

__global__ void k_energy_part(Geo g, int b, int e, const double2 *mm, double gx, double gy, double gz,
                              const double *avg, int pass, double *part);
__global__ void k_energy_final(const double *part, int nb, int n, int pass, double *avg, double *out);
__global__ void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n);

}  // namespace vg
