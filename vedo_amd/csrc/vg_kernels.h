// vg_kernels.h -- kernel argument blocks shared by the kernels and the host.
#pragma once

#include "vg_device.h"

namespace vg {

struct Geo {   // one buffer set of sphere state (see vg_device.h)
    double4 *pos, *vel, *avl;
    double *ori;   // [6][N]
    unsigned *oid;   // sphere rank: position among the spheres in DO-id order
    int *mdl;
};

struct Ell {
    unsigned char *hdr;
    unsigned *ent;    // [2][KI][N]
    double4 *hist;    // [2][KI][N]: shear force xyz, Kn / initial velocity
};

struct Tables {
    const ModelDev *models;
    const LawDev *laws;   // [G][G]
    int n_groups;
};

struct StepArgs {
    int n, nw, ki;
    Geo in, out;
    Ell ell;
    const unsigned *o_cnt;   // candidate-list length per sphere [N]
    const unsigned *o_nbr;   // candidate list, ELL [KO][N]
    int ko;
    Tables t;
    const WallDev *walls;
    int *flags;              // [0..2] relist ring, [3] error bits, [4] max fill
    unsigned long long *relists;
    double dt, gx, gy, gz, skin, relist_half;
    int step;
    int single_group;        // all sphere models share one group: no partner group lookup
    int no_bond;             // no bonded law: surely separated pairs skip the contact geometry
    // record path
    double *rec_imp;         // [6][N] last impulse / angular impulse per sphere (DO order-free, sorted index)
    double *rec;             // [2 halves][KI][13][N] point, dir, to_master, ang_to_master, overlap
};

enum { ERR_INNER_CAP = 1, ERR_OUTER_CAP = 2 };

// Kinematic prescription of a wall: rigid rotation about axis (c, a).
struct WallKin {
    d3 c, a;
    double omega, t_ref;
};

// Cell grid of a rebuild (LeapConsultant.h:103-144).
struct GridDev {
    double xmin, ymin, zmin, zr;
    double safe_v, safe_g;   // dt_up and 0.5*dt_up*dt_up*|g|
    double sf;
    int nx, ny, nz;          // dense dims (reference ncel + 3)
    int perm[3];             // comparator axis priority
};

struct OuterArgs {
    Geo g;
    int n, nw;
    GridDev G;
    const int *cstart, *cend;
    const double *safe;
    const WallDev *walls;
    Tables t;
    int single_group;        // all spheres in one group
    unsigned *count;         // candidates per sphere [N]
    unsigned *nbr;           // ELL [KO][N]
    int ko;
    int *flags;              // [5]: largest candidate count seen
};

struct InnerArgs {
    Geo g, gold;             // new order / previous order
    Ell ell, eold;
    int n, nw, ki;
    const unsigned *perm;    // new index -> previous index
    const unsigned *cnt, *nbr;
    int ko;
    const WallDev *walls;
    double skin;
    int *flags;
};

struct ActiveOut {
    unsigned lo, hi, flags, pad;
    double hist[4];
    double rec[13];
};

template <bool REC, bool SD> __global__ void k_step(StepArgs a);
__global__ void k_walls(WallDev *walls, int nw, const double *assigned, const WallDev *ref, const WallKin *kin,
                        double dt, double t_new, int *flags, int step, double relist_half);
__global__ void k_wall_path_reset(WallDev *walls, int nw);
__global__ void k_bounds(Geo g, int n, unsigned long long *part, unsigned *inv, double *xs, double *ys, double *zs);
__global__ void k_bounds_final(const unsigned long long *part, int nblocks, unsigned long long *red);
__global__ void k_cellkey(Geo g, int n, GridDev G, const unsigned *inv, unsigned long long *keys, unsigned *vals);
__global__ void k_reorder(Geo src, Geo dst, int n, const unsigned *perm, GridDev G, double *safe, int *lin,
                          const double *imp_src, double *imp_dst);
__global__ void k_fill_int(int *p, long long n, int v);
__global__ void k_cellstart(const int *lin, int n, int *cstart, int *cend);
__global__ void k_outer(OuterArgs a);
__global__ void k_count_iact(const unsigned *cnt, const unsigned *nbr, int n, int ko, unsigned long long *out);
__global__ void k_inner_build(InnerArgs a);
__global__ void k_hdr_zero(unsigned char *hdr, int n);
__global__ void k_scatter_state(Geo g, int n, const double *rec_imp, double *out);
__global__ void k_export_active(Geo g, Ell ell, int n, int ki, const unsigned *o_cnt, const unsigned *o_nbr,
                                const WallDev *walls, const unsigned *wall_id, const unsigned *sphere_id,
                                const double *rec, ActiveOut *out, unsigned long long cap,
                                unsigned long long *counter);
__global__ void k_wall_impulse(Ell ell, int n, int ki, const double *rec, double *out);
__global__ void k_assign_spheres(Geo g, int n, const double *buf, double *rec_imp);
__global__ void k_energy_part(Geo g, int n, const double2 *mm, double gx, double gy, double gz, const double *avg,
                              int pass, double *part);
__global__ void k_energy_final(const double *part, int nb, int n, int pass, double *avg, double *out);
__global__ void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n);

}  // namespace vg
