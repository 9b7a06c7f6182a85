/*
 * vedo_oracle.h -- CPU FP64 restatement of VEDO's per-timestep DEM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library (libvedo_hip.so) never
 * links, calls or falls back to it.
 *
 * It follows the reference step in the reference's own operation order so
 * that, compiled without FMA contraction, it reproduces the reference's FP64
 * bits (checked against the compiled reference in oracle/_ref where that
 * builds).  Every function cites the reference file:line it restates.
 *
 * Enumerations use the reference's C++11 values (DOModel.h:26-28):
 *   shape   : 1 Sphere, 6 QuasiCylinder, 7 QuasiPlate
 *   behavior: 1 constrained, 2 fixed, 3 mobile, 4 orbital
 *   scope   : 1 global, 2 local
 */
#ifndef VEDO_ORACLE_H
#define VEDO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { VO_SPHERE = 1, VO_QUASICYLINDER = 6, VO_QUASIPLATE = 7, VO_QUASIPLATE_HOLE = 8 };
enum { VO_CONSTRAINED = 1, VO_FIXED = 2, VO_MOBILE = 3, VO_ORBITAL = 4 };
enum { VO_GLOBAL = 1, VO_LOCAL = 2 };

/* Impact-solver kinds (ModuleList.h:52-56). 0 = no IactModel for the pair. */
enum { VO_LAW_NONE = 0, VO_BSDBF = 1, VO_BSDB2F = 2, VO_BTSDBF = 3, VO_H = 4, VO_HERTZ = 5 };

/* Mechanism slots per law, in the order of the reference constructors:
 *  BSDBF/BSDB2F/BtSDBF (ISwBSDBF.cpp:10-18): NormalStiffness, ShearStiffness,
 *      NormalDampingRatio, ShearDampingRatio, NormalBondStrength,
 *      ShearBondStrength, FrictionCoefficient
 *  H (ISwH.cpp:10-21): HertzNormalStiffness, EquivalentMass,
 *      ShearStiffnessCoefficient1, ShearStiffnessCoefficient2, DampingRatio,
 *      KineticFrictionCoefficient
 *  Hertz (ISwHertz.cpp:11-12): NormalStiffness, NormalDampingCoefficient   */
#define VO_NPARAM 8

typedef struct {
    int shape, behavior, scope, group;
    double density, density_factor;
    double fext[3];
    double attr[3];           /* sphere: r | cylinder: r, h | plate (also with a hole): w, h, l */
    double hole[3];           /* QuasiPlateWithCircularHole: hole radius, x and y offsets */
} vo_model;

typedef struct { int law; double p[VO_NPARAM]; } vo_iact;

/* Per-pair history (ImpactStatus.h:208-222) restricted to what the five
 * solvers read or write. */
typedef struct {
    unsigned char contact, last_contact, bond, last_bond;
    double kn, init_vel;
    double shear[3];
    double point[3], dir[3], to_master[3], ang_to_master[3];
    double overlap;
} vo_status;

typedef struct {
    int n_do, n_models, n_groups;
    const vo_model *models;        /* n_models */
    const vo_iact *iact;           /* n_groups*n_groups, symmetric, first match */
    const int *model_of;           /* n_do */
    const double *pos, *vel, *ox, *oz, *av;   /* 3*n_do, DO id order */
    double dt, t_start, t_stop, t_current;
    double g[3];
    double safety_factor;          /* Constants.h:184 default 1.1 */
    unsigned up_iact, record;
    int consultant;                /* 0 Leap (LeapConsultant), 1 NBS (NBSConsultant) */
    double zoi_lo[3], zoi_hi[3];   /* zone of interest (SystemParameter ZoneOfInterest); an axis
                                      is bounded where lo != hi, none: inactive */
    double pbc_lo[3], pbc_hi[3];   /* periodic boundary (SystemParameter PeriodicBoundaryConditions);
                                      an axis is periodic where lo != hi */
} vo_scene;

typedef struct vo_sim vo_sim;

vo_sim *vo_create(const vo_scene *s, char *err, int errlen);
void vo_destroy(vo_sim *s);

/* Optional history restore before the first step (IactRecordTab restore in
 * SimMediator::Initiate, SimMediator.cpp:690-707).  lo<hi ids. */
int vo_set_history(vo_sim *s, int n, const unsigned *lo, const unsigned *hi, const vo_status *st);

/* One SimMediator::Run() (SimMediator.cpp:714-842).  Returns 1 while
 * t_current < t_stop, like Run(). */
int vo_step(vo_sim *s);

/* Prescribed statuses for kinematic DOs: equivalent of
 * SimMediator::Run(const std::map<_uint_t,DOStatus>&) (SimMediator.cpp:1121-1163):
 * after Response the listed DOs take the given pos/vel/ox/oz/av. */
int vo_step_assigned(vo_sim *s, int n, const unsigned *ids, const double *pos, const double *vel,
                     const double *ox, const double *oz, const double *av);

/* State readback (DOStatus of each DO after the last Response). */
void vo_get_state(const vo_sim *s, double *pos, double *vel, double *ox, double *oz, double *av,
                  double *impact, double *ang_impact);
double vo_time(const vo_sim *s);

/* Candidate list in reference order (IactPairTab), as created interactions. */
int vo_num_interactions(const vo_sim *s);
void vo_get_interactions(const vo_sim *s, unsigned *slave, unsigned *master);
/* Raw candidate pairs (lo,hi) of the last Reset, before interaction creation. */
int vo_num_candidates(const vo_sim *s);
void vo_get_candidates(const vo_sim *s, unsigned *lo, unsigned *hi);

/* Active contacts (ContactInfo.bActive after the last Detect) with their
 * histories keyed by (lo,hi) -- the IactRecordTab payload. Returns count. */
int vo_num_active(const vo_sim *s);
/* The IactRecordTab of the last rebuild or record step: keys in key order. */
int vo_num_records(const vo_sim *s);
void vo_get_records(const vo_sim *s, unsigned *lo, unsigned *hi);
void vo_get_active(const vo_sim *s, unsigned *lo, unsigned *hi, vo_status *st);
/* Current number of DOs (the zone of interest erases mobile DOs that leave
 * it, Consultant::CleanUp) and the scene's original DO id of each. */
int vo_num_do(const vo_sim *s);
void vo_get_ids(const vo_sim *s, unsigned *ids);
/* CleanUp counts so far: DOs erased, IactRecordTab keys moved, moved keys of
 * an erased DO, kept records replaced by a moved one (test instrumentation). */
void vo_cleanup_stats(const vo_sim *s, long long *out);

/* DOWorld::CalculateSystemEnergy (DOWorld.cpp:481-632) of the current state,
 * sequentially over the mobile DOs in DO-id order.  out[16]: potential,
 * translation, rotation, momentum avg (3), angular momentum avg (3) formed as
 * mmi.cross(w), momentum norm, angular momentum norm, v max, v min, w max,
 * w min, number of mobile DOs. */
void vo_system_energy(const vo_sim *s, double *out);

/* Knight analyses of the current state: CalcCoordinateNum::
 * computeAvgCoordinateNum (CalcCoordinateNum.cpp:8-77) and CalcDensity::
 * computeDensity (CalcDensity.cpp:12-72; njr::RandomGenerator seeded with
 * `seed`) over box = xmin xmax ymin ymax zmin zmax; CalculateGranularTemperature
 * (Knight.cpp:1162-1225) of one model over a GSRectangle (centre, Ox, Oz,
 * width, length, height), out[5] = count, |mean v|, T_v, |mean w|, T_w. */
double vo_coordination_number(const vo_sim *s, const double *box);
double vo_packing_density(const vo_sim *s, const double *box, int n_points, int seed);
void vo_granular_temperature(const vo_sim *s, int model, const double *c, const double *ox, const double *oz,
                             const double *size, double *out);

/* Unit-level entry points used by closed-form tests. */
typedef struct { double depth; double n[3]; double p[3]; int active; } vo_contact;
void vo_cd_sphere_sphere(const double *xa, double ra, const double *xb, double rb, vo_contact *c);
void vo_cd_sphere_plate(const double *xa, double ra, const double *xb, const double *oxb,
                        const double *ozb, double w, double h, double l, vo_contact *c);
void vo_cd_sphere_cylinder(const double *xa, double ra, const double *xb, const double *ozb,
                           double rc, double hc, vo_contact *c);
void vo_rotate_around(const double *p, const double *w, double *out);
/* One force-law evaluation on a given contact (A slave, B master).  Writes
 * impulse on A, angular impulses on A and B; updates *st.  Returns 1 if the
 * solver produced an impulse (AddImpact was called). */
int vo_solve(int law, const double *p, double ma, double mb, const vo_contact *c,
             const double *xa, const double *va, const double *wa,
             const double *xb, const double *vb, const double *wb,
             double dt, vo_status *st, double *imp_a, double *ang_a, double *ang_b);

#ifdef __cplusplus
}
#endif
#endif
