/*
 * vedo_gpu.h -- C-ABI of the MI355X DEM step library (libvedo_hip.so).
 *
 * The library replaces the inner work of VEDO's SimMediator::Run()
 * (SimMediator.cpp:714-842) for scenes of spheres against quasi-plates,
 * quasi-plates with a circular hole and quasi-cylinders: the Consultant's
 * candidate list and rebuild cadence (LeapConsultant.cpp:108-195,
 * NBSConsultant.cpp:85-316), the contact
 * detectors (CDSphere2Sphere.cpp:14-60, CDSphere2QuasiPlate.cpp:15-97,
 * CDSphere2QuasiPlateWithCircularHole.cpp:16-134,
 * CDSphere2QuasiCylinder.cpp:16-80), the five impact solvers (ISwBSDBF.cpp,
 * ISwBSDB2F.cpp, ISwBtSDBF.cpp, ISwH.cpp, ISwHertz.cpp), the per-contact
 * ImpactStatus history and its IactRecordTab remap (NBSConsultant.cpp:376-395,
 * SimMediator.cpp:640-712), and DOSphere::Response (DOSphere.cpp:21-48).
 *
 * Conventions:
 *  - every call returns 0 on success and a negative VG_E* code on error; the
 *    message is available from vg_last_error(h).  Reference exit() paths
 *    (missing IactModel mechanism, unknown shape pairing) become errors.
 *  - host buffers are caller-owned; device state is owned by the handle.
 *  - a handle owns one HIP stream and is not thread-safe.
 *  - DO ids are the reference's DOWorld indices (0..n_do-1).
 *  - enumerations use the reference's C++11 values (DOModel.h:26-28).
 */
#ifndef VEDO_GPU_H
#define VEDO_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define VG_OK 0
#define VG_EINVAL (-1)
#define VG_ENOMEM (-2)
#define VG_EHIP (-3)
#define VG_EUNSUPPORTED (-4)
#define VG_ECAPACITY (-5)

/* DOShapeType, DOBehaviorType, DOScopeType (DOModel.h:26-28) */
#define VG_SPHERE 1
#define VG_QUASICYLINDER 6
#define VG_QUASIPLATE 7
#define VG_QUASIPLATE_HOLE 8   /* QuasiPlateWithCircularHole */
#define VG_CONSTRAINED 1
#define VG_FIXED 2
#define VG_MOBILE 3
#define VG_ORBITAL 4
#define VG_GLOBAL 1
#define VG_LOCAL 2

/* Impact solvers (ModuleList.h:52-56).  Mechanism slots per law follow the
 * order in which the reference constructors read them:
 *  ISwBSDBF/ISwBSDB2F/ISwBtSDBF (ISwBSDBF.cpp:10-18): NormalStiffness,
 *    ShearStiffness, NormalDampingRatio, ShearDampingRatio,
 *    NormalBondStrength, ShearBondStrength, FrictionCoefficient
 *  ISwH (ISwH.cpp:10-21): HertzNormalStiffness, EquivalentMass,
 *    ShearStiffnessCoefficient1, ShearStiffnessCoefficient2, DampingRatio,
 *    KineticFrictionCoefficient
 *  ISwHertz (ISwHertz.cpp:11-12): NormalStiffness, NormalDampingCoefficient */
#define VG_LAW_NONE 0
#define VG_BSDBF 1
#define VG_BSDB2F 2
#define VG_BTSDBF 3
#define VG_H 4
#define VG_HERTZ 5
#define VG_NPARAM 8

/* DOModel (DOModel.h:132-137) after DOModel::Analysis (DOModel.cpp:17-149). */
typedef struct {
    int shape, behavior, scope, group;
    double density, density_factor;
    double fext[3];
    /* Sphere: radius | QuasiCylinder: radius, height | QuasiPlate: width,
     * height, length | QuasiPlateWithCircularHole: width, height, length,
     * hole radius, hole x / y offset (DOModel.cpp:90-120) */
    double attr[6];
} vg_model;

/* One entry of the dense group x group IactModel table (IactModel.cpp:16-30,
 * DOWorld.cpp:297-302: symmetric, first match). */
typedef struct {
    int law;
    double p[VG_NPARAM];
} vg_iact;

/* A DOWorld in structure-of-arrays form (replaces the DOWorld/SystemParameter
 * the Consultant reads, DOWorld.h:32, SystemParameter.h:17-25). */
typedef struct {
    int n_do, n_models, n_groups;
    const vg_model *models;
    const vg_iact *iact;       /* n_groups*n_groups */
    const int *model_of;       /* n_do */
    const double *pos, *vel, *ox, *oz, *av;   /* 3*n_do each */
    double dt, t_start, t_stop, t_current;
    double g[3];
    double safety_factor;      /* aries::Constants SafetyFactor (Constants.h:184) */
    unsigned up_iact;          /* candidate-list rebuild period (anne's UpIact) */
    unsigned record;           /* record period (anne's Record) */
    int consultant;            /* candidate list: 0 LeapConsultant's, 1 NBSConsultant's
                                  (NBSConsultant.cpp:85-374; one-scene handles, no zone of interest) */
    double inner_skin;         /* <= 0: automatic */
    /* SystemParameter ZoneOfInterest (SystemParameter.h:275): an axis is
     * bounded where lower != upper (Boundary::Correct, Boundary.cpp:79-129);
     * with a bounded axis, mobile DOs outside are erased at construction and
     * at every candidate-list rebuild (Consultant::CleanUp, Consultant.cpp:
     * 264-319) and the remaining DO ids close up. */
    double zoi_lo[3], zoi_hi[3];
    /* SystemParameter PeriodicBoundaryConditions (SystemParameter.h:276): an
     * axis is periodic where lower != upper.  Positions are taken back into
     * [lower, upper) after Response (Boundary::EnforceBoundaryConditions,
     * SimMediator.cpp:746), contact, lever-arm and candidate distances are
     * minimum images (Boundary::DifferenceBoundaryConditions).  Only with
     * vg_create (one scene), fixed walls inside the range and no prescribed
     * statuses or wall rotation. */
    double pbc_lo[3], pbc_hi[3];
} vg_scene_desc;

/* ImpactStatus payload of one contact (ImpactStatus.h:208-222). */
typedef struct {
    unsigned char contact, last_contact, bond, last_bond;
    double kn, init_vel;
    double shear[3];
    double point[3], dir[3], to_master[3], ang_to_master[3];
    double overlap;
} vg_status;

typedef struct vg_scene vg_scene;

/* Create a scene on the current HIP device: builds the candidate list
 * (Consultant::Reset) and the interactions (SimMediator::Initiate).
 * Replaces `new LeapConsultant(...)` + `SimMediator sm(...)` (Anne.cpp:90-118). */
int vg_create(const vg_scene_desc *desc, vg_scene **out);
int vg_destroy(vg_scene *h);
const char *vg_last_error(const vg_scene *h);
/* Error text of the last failed vg_create (no handle exists then). */
const char *vg_create_error(void);

/* Restore histories before the first step (IactRecordTab restore,
 * SimMediator.cpp:690-707). lo < hi are DO ids. */
int vg_set_history(vg_scene *h, int n, const unsigned *lo, const unsigned *hi, const vg_status *st);

/* Enqueue nsteps of SimMediator::Run() (SimMediator.cpp:714-842) on the
 * handle's stream.  Asynchronous; vg_sync waits.  Rebuilds every up_iact
 * steps as the Consultant's NextStep does (LeapConsultant.cpp:49-97). */
int vg_step(vg_scene *h, int nsteps);
int vg_sync(vg_scene *h);

/* One step with assigned statuses, the equivalent of
 * SimMediator::Run(const std::map<_uint_t,DOStatus>&) (SimMediator.cpp:
 * 1121-1163, 843-889): after Response the listed DOs (spheres or walls) take
 * the given status, with zero stored impulses, before contact detection and
 * the rebuild.  Arrays hold 3 doubles per listed DO.  Synchronous. */
int vg_step_assigned(vg_scene *h, int n, const unsigned *ids, const double *pos, const double *vel,
                     const double *ox, const double *oz, const double *av);

/* Prescribe a rigid rotation of walls about an axis, evaluated on the device
 * every step (a kinematic drum), instead of host-assigned statuses.  The
 * listed walls' current statuses are the t = t_ref configuration. */
int vg_set_wall_rotation(vg_scene *h, int n, const unsigned *ids, const double *axis_point,
                         const double *axis_dir, double omega, double t_ref);

/* DOStatus of every DO in DO-id order, as DOWorld::UpdateDOStatus sees it
 * (DOWorld.cpp:304-318); impact/ang_impact are the impulses of the last
 * Response.  Any pointer may be NULL. */
int vg_get_state(vg_scene *h, double *pos, double *vel, double *ox, double *oz, double *av,
                 double *impact, double *ang_impact);
double vg_time(const vg_scene *h);

/* Current number of DOs of a scene (slot 0 of a one-scene handle): the zone
 * of interest erases DOs, and vg_get_state / vg_scene_get_state return the
 * remaining ones in their (closed-up) DO-id order. */
int vg_scene_num_do(vg_scene *h, int scene, int *n_do);

/* DOModel index (into the description's models) of each remaining DO of a
 * scene, in the DO-id order vg_scene_num_do counts: the DOName each DOStatus
 * carries (DOStatus::GetDOName, src/vedo/framework/DOStatus.h:78), so that a
 * state taken after the zone of interest erased DOs can seed a new scene. */
int vg_scene_models(vg_scene *h, int scene, int *model_of);

/* Number of interactions in the candidate list (Consultant::GetIactNum). */
int vg_num_interactions(vg_scene *h, long long *n);

/* Neighbour-list skin (distance beyond contact within which candidates are
 * walked every step).  Any value gives the same results; it trades per-step
 * list length against relist frequency.  No reference counterpart: the
 * reference walks its whole candidate list every step. */
int vg_set_skin(vg_scene *h, double skin);
double vg_get_skin(const vg_scene *h);
/* Skin per scene from its speed: at every rebuild each scene's skin becomes
 * 2.2 times the path its fastest sphere can cover in a rebuild period (speed
 * and gravity), clamped to [floor, cap], so that relists between rebuilds
 * stay rare while slow scenes keep short lists.  Results do not depend on it.
 * vg_set_skin switches back to one fixed skin.  No reference counterpart. */
int vg_set_skin_auto(vg_scene *h, double floor, double cap);
/* Arithmetic of the pair solve.  VG_PRECISION_FP64 (the default) is the
 * reference's: every operation in FP64, in its order, so results are the
 * reference's bits.  VG_PRECISION_MIXED keeps in FP64 the contact depth (so
 * the active set and every ImpactStatus flag are the FP64 ones), the stored
 * histories, the impulse sums and the integration, and forms the contact
 * normal, lever arms, relative velocity, forces and torques of each pair in
 * FP32 (per-pair force error of a few FP32 rounding units).  Takes effect
 * at the next step.  No reference counterpart: the reference is FP64 only. */
#define VG_PRECISION_FP64 0
#define VG_PRECISION_MIXED 1
int vg_set_precision(vg_scene *h, int precision);
int vg_get_precision(const vg_scene *h);

/* Active contacts (ContactInfo.bActive of the last detect) with their
 * ImpactStatus, keyed (lo, hi) -- the IactRecordTab payload that
 * RebuildIactRecordTab saves (NBSConsultant.cpp:376-395).  Call with
 * st == NULL to get the count in *n; otherwise *n is the capacity on input. */
int vg_get_active(vg_scene *h, long long *n, unsigned *lo, unsigned *hi, vg_status *st);

/* Timing on the handle's stream.  After vg_kernel_timing(h, k) the next k
 * step-kernel launches are bracketed by device events and vg_step_timing
 * returns their summed duration and count; otherwise it returns the duration
 * of the last vg_step call (all its kernels) and its step count. */
int vg_kernel_timing(vg_scene *h, int max_launches);
int vg_step_timing(vg_scene *h, double *ms, int *launches);
/* No counterpart (measurement): after vg_kernel_timing(h, k), the candidate-
 * list rebuilds of the following steps (every UpIact-th step's
 * LeapConsultant::Reset + history carry, host work included) are bracketed by
 * device events too; their summed duration and count. */
int vg_rebuild_timing(vg_scene *h, double *ms, int *rebuilds);
/* No counterpart (diagnostic): checks on the device that every index the
 * step, wall and rebuild kernels take from the handle's lists is in range
 * (each sphere's model and scene; its neighbour-list and candidate-list
 * counts and entries: placed spheres of its own scene or wall slots of its
 * scene's range; the history slots of its contacts: within the list
 * capacity, none shared; its constrained-wall and pair slots).  counts[0..7]:
 * violations by kind (0 model, 1 scene, 2 neighbour count, 3 neighbour
 * entry, 4 history slot, 5 candidate count, 6 candidate entry, 7
 * constrained-wall slot);
 * counts[8..10]: the first violation's kind + 1, sphere slot and entry. */
int vg_check_lists(vg_scene *h, long long counts[11]);
/* Counters since vg_create: rebuilds, relists, max neighbour-list fill, the
 * list capacity, and contact_steps, the sum over steps of the number of
 * sphere-sphere and sphere-wall pairs left in contact by the step's solve.
 * Any pointer may be NULL. */
int vg_stats(vg_scene *h, long long *rebuilds, long long *relists, int *max_fill, int *capacity,
             long long *contact_steps);

/* DOWorld::CalculateSystemEnergy (DOWorld.cpp:481-632) over the mobile DOs,
 * computed on the device from the current state.  out[16]: potential,
 * translational and rotational energy, average momentum (3), average
 * "angular momentum" as the reference forms it, mmi.cross(w) (3), momentum
 * norm, angular momentum norm, max |v|, min |v|, max |w|, min |w|, number of
 * mobile DOs.  Sums are reduced in a fixed tree order. */
int vg_system_energy(vg_scene *h, double *out);

/* ---- Batched and served scenes ----
 * A batch handle holds many independent scenes in one packed device layout
 * and steps all of them with one kernel launch per step.  Each scene keeps
 * its own DOWorld (models, IactModel table, walls, dt, gravity), grid and
 * candidate order, so its result is the same alone and in a batch.  The
 * scenes share UpIact; a scene is placed on the device, or leaves it, at a
 * rebuild boundary (every UpIact batch steps), so every scene rebuilds at its
 * own steps 0, UpIact, 2 UpIact, ... as it would alone.  No reference
 * counterpart: the reference runs one DOWorld per process (Anne.cpp). */
#define VG_SCENE_FREE 0
#define VG_SCENE_QUEUED 1     /* admitted, placed at the next boundary */
#define VG_SCENE_RUNNING 2
#define VG_SCENE_FINISHED 3   /* reached its step count; state held until the next boundary */
#define VG_SCENE_RETIRING 4   /* leaves the device at the next boundary */
#define VG_SCENE_DONE 5       /* left the device; its final state is kept on the host */
#define VG_SCENE_FAILED 6     /* blew up (a sphere crossing more than the scene's largest radius in one
                                 step) or outgrew the candidate capacity, and left the batch */

typedef struct {
    long long max_spheres;     /* sphere capacity of the device layout */
    int max_walls;             /* wall capacity */
    int max_scenes;            /* scene slots */
    int neighbour_capacity;    /* neighbour-list entries per sphere to start with (<= 127; 0: 32); rebuilds
                                  grow it, up to 127, when a sphere has more partners within the skin */
    unsigned up_iact;          /* shared UpIact */
} vg_batch_desc;

int vg_batch_create(const vg_batch_desc *bd, vg_scene **out);
/* Queue a scene for `steps` steps (< 0: unbounded); *scene is its slot. */
int vg_admit(vg_scene *h, const vg_scene_desc *desc, long long steps, int *scene);
/* Drop a scene (at the next boundary if it is on the device) or forget a
 * kept result. */
int vg_retire(vg_scene *h, int scene);
int vg_scene_status(vg_scene *h, int scene, int *state, long long *steps_done, double *t);
/* DOStatus of the scene's DOs in its DO-id order (kept on the host once the
 * scene is done). */
int vg_scene_get_state(vg_scene *h, int scene, double *pos, double *vel, double *ox, double *oz, double *av);
int vg_scene_get_active(vg_scene *h, int scene, long long *n, unsigned *lo, unsigned *hi, vg_status *st);
int vg_scene_energy(vg_scene *h, int scene, double *out);

/* Knight analyses (the knight tool, src/vedo/knight) of a scene's current
 * state on the device (slot 0 for a one-scene handle).  box = xmin, xmax,
 * ymin, ymax, zmin, zmax; a sphere belongs to it when its centre lies within
 * the box widened by its radius (strict comparisons).
 *
 * CalcCoordinateNum::computeAvgCoordinateNum
 * (src/vedo/knight/CalcCoordinateNum.cpp:8-77): the mean number of spheres
 * whose centre distance is at most the sum of the radii, over the box's
 * spheres (NaN for none).  Touching pairs are read from the candidate list:
 * every pair of sphere groups needs an IactModel (VG_EUNSUPPORTED otherwise). */
int vg_scene_coordination_number(vg_scene *h, int scene, const double box[6], double *avg);
/* CalcDensity::computeDensity (src/vedo/knight/CalcDensity.cpp:12-72): the
 * percentage of n_points random points of the box (njr::RandomGenerator,
 * src/njr/RandomGenerator.cpp, seeded with `seed`; the reference's first
 * generator in a process uses 999 and RandomHitTestNum is 20000) that lie in
 * a sphere of the box. */
int vg_scene_packing_density(vg_scene *h, int scene, const double box[6], int n_points, int seed, double *percent);
/* vg_scene_packing_density without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * around HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, same result bit for bit; the sample points and
 * the hit counter live in a fixed per-GPU scratch array of the library
 * (VG_DENSITY_NOALLOC_MAX_POINTS points, 1.5 MiB, held until the process
 * ends), so n_points above that maximum is rejected (VG_EINVAL).  Every call
 * on one GPU shares that GPU's scratch: calls that overlap on one GPU (two
 * threads, two handles) race and must be serialised by the caller. */
#define VG_DENSITY_NOALLOC_MAX_POINTS 65536
int vg_scene_packing_density_noalloc(vg_scene *h, int scene, const double box[6], int n_points, int seed,
                                     double *percent);
/* CalculateGranularTemperature (src/vedo/knight/Knight.cpp:1162-1225) of the
 * spheres of DOModel `model` (the description's index) inside a GSRectangle
 * (src/vedo/common/GSRectangle.cpp:15-30: centre, axes Ox and Oz, width along
 * Ox, length along Oz x Ox, height along Oz).  out[5] = count, |mean v|,
 * T_v, |mean w|, T_w.  The sums run in a fixed order of their own, not the
 * reference's DO-id order: results agree to rounding (1e-12 relative). */
int vg_scene_granular_temperature(vg_scene *h, int scene, int model, const double center[3], const double ox[3],
                                  const double oz[3], const double size[3], double out[5]);
/* vg_set_wall_rotation for one scene of a batch (also before placement). */
int vg_scene_set_wall_rotation(vg_scene *h, int scene, int n, const unsigned *ids, const double *axis_point,
                               const double *axis_dir, double omega, double t_ref);
int vg_batch_info(vg_scene *h, long long *spheres, int *scenes_on_device, int *scenes_queued,
                  unsigned long long *steps);
/* No counterpart (batch handles): step whole rebuild periods as captured HIP
 * graphs (on = 1, the default) or launch every step from the host (on = 0).
 * A period is replayed from a graph when it starts at a rebuild, no scene
 * reaches its step count inside it and no wall is constrained; results are
 * the same either way.  vg_graph_replays: periods replayed so far.  The
 * environment variable VG_GRAPHS=0 makes off the default of new batch
 * handles (for tools that trace each kernel launch). */
int vg_set_graphs(vg_scene *h, int on);
int vg_graph_replays(vg_scene *h, long long *replays);

/* Streaming copy of `bytes` bytes, `iters` times, on the current device;
 * returns the achieved read+write bandwidth in GB/s.  Diagnostics: measures
 * the achievable HBM bandwidth and calibrates traffic counters.  No reference
 * counterpart. */
int vg_copy_bench(size_t bytes, int iters, double *gbps);

#ifdef __cplusplus
}
#endif
#endif
