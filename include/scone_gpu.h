/*
 * scone_gpu.h -- C ABI of the MI355X multigroup transport engine (libscone_amd.so).
 *
 * The engine replaces the per-history OpenMP loop of SCONE's physics packages
 * (PhysicsPackages/eigenPhysicsPackage_class.f90:211-252, fixedSourcePhysicsPackage_class.f90:
 * 198-261) by a bank-level call.  Everything the loop reads is passed once as flat tables
 * (geometry, multigroup data, collision-estimator clerks); banks cross the boundary as arrays
 * of scn_particle_state, the bind(C) image of particleStateData (ParticleObjects/particle_class.f90:50-65).
 *
 * Conventions
 *  - All indices the reference keeps 1-based stay 1-based (material, universe, local cell,
 *    group, broodID); array offsets into the flat tables are 0-based.
 *  - Every entry point returns 0 on success or an SCN_ERR_* code; scn_last_error() then holds
 *    a message naming the reference check it mirrors (fatalError sites).
 *  - One engine handle is used by one host thread at a time.  Several handles may coexist.
 */
#ifndef SCONE_GPU_H
#define SCONE_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes -------------------------------------------------------------------------- */
enum {
    SCN_OK = 0,
    SCN_ERR_ARG = 1,         /* invalid argument or table                                       */
    SCN_ERR_HIP = 2,         /* HIP runtime failure                                             */
    SCN_ERR_OVERFLOW = 3,    /* bank / site buffer overflow (particleDungeon_class.f90:170-174) */
    SCN_ERR_MATERIAL = 4,    /* UNDEF / OVERLAP material or invalid material index              */
    SCN_ERR_NAN = 5,         /* NaN distance (transportOperatorST_class.f90:86-91)              */
    SCN_ERR_MAJORANT = 6,    /* zero majorant (transportOperatorDT_class.f90:60)                */
    SCN_ERR_NORM = 7,        /* population normalisation failed (particleDungeon_class.f90:597) */
    SCN_ERR_SAMPLING = 8,    /* CDF scan fell through (multiScatterMG/fissionMG sampleOut)      */
    SCN_ERR_UNSUPPORTED = 9, /* feature outside the engine's scope                              */
    SCN_ERR_NO_DEVICE = 10,  /* no GPU visible                                                  */
    SCN_ERR_BUSY = 11        /* batch engine: no free slot or HBM budget for the request now    */
};

/* ---- reference constants ------------------------------------------------------------------ */
#define SCN_MAX_NEST 8            /* deepest universe nesting the engine accepts (reference: 12) */
#define SCN_OUTSIDE_MAT 0
#define SCN_VOID_MAT 2147483647
#define SCN_UNDEF_MAT 2147483646
#define SCN_OVERLAP_MAT 2147483645
#define SCN_BC_VACUUM 0
#define SCN_BC_REFLECTIVE 1
#define SCN_BC_PERIODIC 2

enum { SCN_TRANSPORT_ST = 1, SCN_TRANSPORT_DT = 2, SCN_TRANSPORT_HT = 3 };
/* neutronMGstd (neutronMGstd_class.f90); neutronMGimp (neutronMGimp_class.f90), whose fission
 * sites keep the colliding particle's collision count instead of restarting at 0 */
enum { SCN_COLL_STD = 0, SCN_COLL_IMP = 1 };
enum { SCN_UNI_ROOT = 1, SCN_UNI_LAT = 2, SCN_UNI_PIN = 3, SCN_UNI_CELL = 4 };
/* unionCell records and tokens (operators above every surface index) */
enum { SCN_UCELL = -1 };
enum { SCN_UCELL_UNION = 1073741824, SCN_UCELL_INTERSECT = 1073741825, SCN_UCELL_OPEN = 1073741826,
       SCN_UCELL_CLOSE = 1073741827, SCN_UCELL_COMPLEM = 1073741828 };
enum { SCN_SURF_APLANE = 1, SCN_SURF_PLANE = 2, SCN_SURF_SPHERE = 3, SCN_SURF_CYLINDER = 4,
       SCN_SURF_QUADRIC = 5, SCN_SURF_BOX = 6, SCN_SURF_TRUNC_CYLINDER = 7,
       SCN_SURF_TRUNC_CONE = 8, SCN_SURF_WEDGE = 9 };
enum { SCN_MAP_SPACE = 1, SCN_MAP_MATERIAL = 2, SCN_MAP_ENERGY = 3, SCN_MAP_CELL = 4 };
enum { SCN_GRID_LIN = 1, SCN_GRID_UNSTRUCT = 3 };
/* macroscopic MT codes (SharedModules/endfConstants.f90); 0 selects fluxResponse */
enum {
    SCN_MT_FLUX = 0, SCN_MT_TOTAL = -1, SCN_MT_DISAPPEARANCE = -2, SCN_MT_ESCATTER = -3,
    SCN_MT_IESCATTER = -4, SCN_MT_FISSION = -6, SCN_MT_NUFISSION = -7, SCN_MT_KAPPA = -80,
    SCN_MT_ALLSCATTER = -20, SCN_MT_ABSORPTION = -21, SCN_MT_NONELASTIC = -22
};

/* ---- bank record: particleStateData (particle_class.f90:50-65, MPI layout mpi_func.f90:95-113) */
typedef struct scn_particle_state {
    double wgt;
    double r[3];
    double dir[3];
    double E;
    int32_t G;
    int32_t isMG;
    int32_t type;
    int32_t pad0;
    double time;
    double lambda;
    int32_t matIdx;
    int32_t cellIdx;
    int32_t uniqueID;
    int32_t collisionN;
    int32_t broodID;
    int32_t pad1;
} scn_particle_state;

/* ---- geometry ------------------------------------------------------------------------------
 * Border surfaces: box (Geometry/Surfaces/CompositeSurfaces/box_class.f90) with n_axes = 3, or
 * x/y/zSquareCylinder (squareCylinder_class.f90) with n_axes = 2.  origin/halfwidth are stored
 * per slot; axes[slot] is the coordinate (0 = x, 1 = y, 2 = z) of that slot.                  */
typedef struct scn_box {
    int32_t n_axes;
    int32_t axes[3];
    int32_t bc[6];           /* -x +x -y +y -z +z */
    double origin[3];
    double halfwidth[3];
} scn_box;

/* Surfaces of cells and of a non-box root border (Geometry/Surfaces).  tol is surfTol: SURF_TOL,
 * or what the type's init sets (sphere, cylinder 2*r*SURF_TOL; quadric 2*max|coeff|*SURF_TOL).
 *  APLANE         x/y/zPlane (aPlane_class.f90): axis, c[0] = a0
 *  PLANE          plane (plane_class.f90): c[0..2] normal, c[3] offset, both divided by
 *                 norm2(normal) as init does
 *  SPHERE         sphere (sphere_class.f90): c[0..2] origin, c[3] radius, c[4] radius**2
 *  CYLINDER       x/y/zCylinder (cylinder_class.f90): axis, c[0..2] origin, c[3] radius, c[4] radius**2
 *  QUADRIC        quadric (quadric_class.f90): c[0..9] the ten coefficients
 *  BOX            box or x/y/zSquareCylinder: axis = index into boxes
 *  TRUNC_CYLINDER x/y/zTruncCylinder (truncCylinder_class.f90): axis, c[0..2] origin, c[3] radius,
 *                 c[4] halfwidth
 *  TRUNC_CONE     x/y/zTruncCone (truncCone_class.f90): axis, c[0..2] vertex, c[3] tan(angle)**2,
 *                 c[4] hMin, c[5] hMax; tol = SURF_TOL * (|hMin| + |hMax|)/2 * tan(angle)
 *  WEDGE          x/y/zWedge (wedge_class.f90): axis, c[0..2] origin, c[3] halfwidth, c[4] altitude,
 *                 c[5..6] normal1, c[7..8] normal2, c[9..10] normal3 (in-plane components, as
 *                 build computes them from opening and rotation); tol = SURF_TOL * halfwidth    */
typedef struct scn_surface {
    int32_t type;
    int32_t axis;
    double tol;
    double c[12];
} scn_surface;

/* Universes (Geometry/Universes/{root,lat,pin,cell}Universe_class.f90).  uniIdx is the 1-based position in the table.
 *  ROOT: rootUniverse_class.f90 -- local cell 1 = inside border, 2 = outside.  A box or
 *        squareCylinder border is boxes[border] (border_surf = -1); any other border surface is
 *        surfaces[border_surf] (vacuum boundary only, as surface_inter.f90 setBC requires).
 *  LAT : latUniverse_class.f90  -- size/pitch after the 2D fix-up (size[2] = 1, pitch[2] = 2*INF),
 *        corner, a_bar and out_local as init computes them; offset_mode 1 = offset every cell,
 *        0 = none, 2 = per-cell flags at int_data[offset_map + localID - 1].
 *  PIN : pinUniverse_class.f90  -- annuli sorted ascending, the last one of radius INF*1.1;
 *        real_data[annuli + i] = radius_i, real_data[annuli + n_annuli + i] = radius_i**2.
 *  CELL: cellUniverse_class.f90 -- n_cells cells in definition order (local IDs 1..n_cells),
 *        local IDs n_cells+1 / n_cells+2 are the UNDEF / OVERLAP cells.
 *        int_data[cells + i] is the offset in int_data of local cell i+1's record and
 *        int_data[cells + n_cells + i] its cellShelf index (1-based, geometry cells in definition
 *        order).  A simpleCell (simpleCell_class.f90) record is [n, s_1 .. s_n]:
 *        s_k = +-(1-based index into surfaces), + for the positive halfspace.  A unionCell
 *        (unionCell_class.f90) record is [SCN_UCELL, simple, m, t_1 .. t_m]: the evaluation tokens
 *        after deMorgan, a surface token being +-(1-based surface index) and an operator one of
 *        SCN_UCELL_UNION .. SCN_UCELL_COMPLEM; simple = 1 when no union is left (inside then tests
 *        the surfaces in order, as the reference does).
 * Every type: origin, and rotMat (universe_inter.f90 setTransform, rotationMatrix of
 * genericProcedures.f90) when rotated = 1, stored row-major: rot[3*(i-1) + j-1] = rotMat(i,j).
 * global = 1: the universe is entered from the global position ('global' of setupBase).
 * check_overlap = 1 (CELL): every cell is tested and a point inside several is in the OVERLAP
 * cell (findCellOverlap); otherwise the first cell in definition order that holds it. */
typedef struct scn_universe {
    int32_t type;
    int32_t n_local;
    double origin[3];
    int32_t border;          /* ROOT: index into boxes                       */
    int32_t size[3];         /* LAT */
    int32_t out_local;
    int32_t offset_mode;
    int32_t offset_map;
    double pitch[3];
    double corner[3];
    double a_bar[3];
    double outline_hw[3];
    int32_t n_annuli;        /* PIN */
    int32_t annuli;
    int32_t n_cells;         /* CELL */
    int32_t cells;
    int32_t border_surf;     /* ROOT */
    int32_t rotated;
    int32_t global;
    int32_t check_overlap;   /* CELL */
    double rot[9];
} scn_universe;

/* ---- multigroup data (NuclearData/mgNeutronData/baseMgNeutron, Reactions/reactionMG) -----------
 * xs[(m-1)*nG*6 + (g-1)*6 + k], k = TOTAL, IESCATTER, CAPTURE, FISSION, NU_FISSION, KAPPA_FISSION
 * as baseMgNeutronMaterial%data holds them (fission columns zero for non-fissile materials).
 * P0/prod[(m-1)*nG*nG + (gin-1)*nG + (gout-1)] is column gin of P0(G_out, G_in) (contiguous scan).
 * chi[(m-1)*nG + g-1] is the normalised fission spectrum.                                     */
typedef struct scn_mg_data {
    int32_t n_mat;
    int32_t n_groups;
    const double* xs;
    const double* P0;
    const double* prod;
    const double* chi;
    const int32_t* fissile;
    const double* majorant;  /* [nG] baseMgNeutronDatabase_class.f90:458-503 */
    double collision_xs;     /* 1/avgDist, 0 when absent */
    const double* P1;        /* PN P1: [mat][G_in][G_out] 3*P1/P0 (multiScatterP1MG buildFromDict), NULL for P0 */
} scn_mg_data;

/* ---- tallies: collisionClerk (Tallies/TallyClerks/collisionClerk_class.f90) ----------------- */
typedef struct scn_map1d {
    int32_t type;            /* SCN_MAP_SPACE / _MATERIAL / _ENERGY (multigroup: always bin 0,
                                energyMap_class.f90 map) / _CELL (cellMap: int_data[data + c] = bin
                                of cellShelf index c in 0..axis, c = 0 outside cell universes) */
    int32_t axis;            /* SPACE: 0,1,2 */
    int32_t nbins;
    int32_t grid_type;       /* SPACE: SCN_GRID_LIN (real_data[data] = min, [data+1] = step)
                                       SCN_GRID_UNSTRUCT (real_data[data .. data+nbins]) */
    int32_t data;            /* MATERIAL: int_data[data + matIdx] = bin for matIdx in 0..n_mat,
                                int_data[data + n_mat + 1] = bin for any other index */
} scn_map1d;

/* Clerks (Tallies/TallyClerks): kind SCN_CLERK_COLLISION = collisionClerk (collisionClerk_class.f90,
 * scores w/Sigma at every collision), SCN_CLERK_TRACK = trackClerk (trackClerk_class.f90, scores
 * w*L for every surface-tracking path of length L, with map and response evaluated at the
 * pre-path state; delta-tracking flights report no path). */
enum { SCN_CLERK_COLLISION = 0, SCN_CLERK_TRACK = 1 };
typedef struct scn_clerk {
    int32_t gtop, glow;      /* energyFilter window; gtop = 0 means no filter */
    int32_t n_maps;          /* 0 (single bin) .. 3 (multiMap) */
    int32_t handle_virtual;
    scn_map1d maps[3];
    int32_t multi[3];
    int32_t width;           /* number of responses */
    int32_t resp[8];         /* SCN_MT_* */
    int32_t kind;            /* SCN_CLERK_* */
    int64_t mem_addr;        /* 0-based offset of this clerk's first bin in the score memory */
    int64_t size;            /* width * product(nbins) */
} scn_clerk;

typedef struct scn_problem {
    int32_t transport;       /* SCN_TRANSPORT_ST / _DT / _HT (transportOperatorFactory_func.f90:36-67) */
    int32_t st_cache;        /* transportOperatorST 'cache' flag (default 1) */
    int32_t n_universes;
    int32_t root_uni;        /* uniIdx of the root universe */
    const scn_universe* universes;
    int32_t n_boxes;
    int32_t border_box;
    const scn_box* boxes;
    int32_t graph_size;      /* geomGraph: fill >= 0 material (0 outside), < 0 -uniIdx */
    const int32_t* graph_fill;
    const int32_t* graph_id; /* uniqueID (material cells) or rootID of the filling universe */
    int32_t n_real;
    const double* real_data;
    int32_t n_int;
    const int32_t* int_data;
    scn_mg_data mg;
    int32_t n_clerks;
    const scn_clerk* clerks;
    int64_t n_bins;          /* size of the collision-clerk score memory */
    int64_t norm_bin;        /* tallyAdmin 'norm': 0-based BIN address of the norm clerk, -1 none */
    double norm_value;       /* tallyAdmin 'normVal' */
    double ht_cutoff;        /* transportOperatorHT 'cutoff' (default 0.9): DT when sigmaT/majorant > 1 - cutoff */
    int32_t n_surfaces;      /* surfaces of cells and of a non-box root border */
    const scn_surface* surfaces;
    int32_t collision;       /* SCN_COLL_STD / SCN_COLL_IMP (collisionProcessorFactory_func.f90) */
    int32_t max_split;       /* neutronMGimp 'maxSplit' (default 1000) */
    /* neutronMGimp weight windows (weightWindowsField_class.f90; fixed-source runs): ww_n bins
     * located by ww_map (n_maps/maps/multi as a clerk's), lower and upper weights per bin, and
     * constSurvival.  After a collision a particle heavier than the bin's upper weight is split
     * into ceiling(w/upper) (maxSplit in all), the copies going to the history's buffer; one
     * lighter than the lower weight survives Russian roulette with weight lower*constSurvival.
     * ww_n = 0: no weight windows. */
    int32_t ww_n;
    int32_t fixed_point_tally; /* 1: collision/track clerk BINs accumulate as 128-bit fixed point (each score
                                  v * 2^64 truncated, added as four 32-bit limbs into 64-bit words): the
                                  sums are the same bits whatever the order the scores arrive in, so
                                  runs and batched/plain solves give identical results.  0: FP64 atomics
                                  (faster; sums equal to rounding of summation order). */
    const scn_clerk* ww_map;
    const double* ww_lower;
    const double* ww_upper;
    double ww_survival;
} scn_problem;

/* ---- eigenvalue run ------------------------------------------------------------------------ */
typedef struct scn_eigen_settings {
    int32_t pop;             /* totalPop */
    int32_t n_inactive;
    int32_t n_active;
    int32_t inactive_tally;  /* 1: the clerks also score during inactive cycles (own memory) */
    int64_t seed;            /* pRNG initial state (eigenPhysicsPackage_class.f90:492-495) */
    double keff0;            /* keff_0, default 1.0 */
    int32_t events_per_launch; /* 0: autotuned schedule */
    int32_t block_size;        /* 0: autotuned */
    int32_t cycle_graph;       /* 1: every size and scalar of a cycle stays on the device and the cycles
                                  are replayed from a captured hipGraph (one graph per phase); the
                                  same histories, banks and k history as 0.  Needs events_per_launch 0. */
    int32_t refill;            /* idle lanes a wave gathers before it starts new histories (1..64; 0: default);
                                  the histories do not change */
    double select_window;      /* cycle_graph: half-width, in standard deviations of the order statistic, of
                                  the window of draws gathered to find normSize_Repr's threshold (0: 8).
                                  The threshold is exact for any value: a draw outside the window only
                                  moves the search to a pass over all draws. */
} scn_eigen_settings;

typedef struct scn_eigen_result {
    double k_analog[2];      /* inactive keffAnalogClerk K_EST mean, sigma */
    double k_implicit[2];    /* active keffImplicitClerk K_EFF mean, sigma */
    double imp[4][2];        /* IMP_PROD, SCATTER_PROD, IMP_ABS, ANA_LEAK mean, sigma (active) */
    double* k_history;       /* caller buffer [n_inactive + n_active] or NULL: k used next cycle */
    double* cycle_time;      /* caller buffer [n_inactive + n_active] or NULL: wall seconds per cycle */
    double* bins_active;     /* caller buffers [n_bins * 2] (mean, sigma) or NULL */
    double* bins_inactive;
    double transport_time;   /* seconds, same span as SCONE's Total_Transport_Time */
    int64_t n_collisions;    /* real collisions */
    int64_t n_events;        /* flights + collisions processed */
    int64_t n_histories;
} scn_eigen_result;

/* ---- fixed-source problems ------------------------------------------------------------------
 * Sources of ParticleObjects/Source, sampled on the device with the particle's own stream:
 *   POINT    pointSource_class.f90:91-270 through configSource_inter.f90:129-141: position r,
 *            direction isotropic (mu = 2u-1, phi = 2 pi u, rotateVector([1,0,0])) or `dir`, then
 *            the group: G, or a scan of probG with one draw.  wgt 1, time 0.
 *   MATERIAL materialSource_class.f90:150-231: per attempt three position draws in
 *            [bottom, top) and one time draw; rejected while outside the geometry or in another
 *            material (VOID/UNDEF/OVERLAP is an error); 200 attempts; then mu, phi; group G.
 * `dir` and `probG` are given as written in the input: the library normalises them as
 * pointSource%init does (dir / norm2(dir), probG / sum(probG)).                                  */
enum { SCN_SRC_POINT = 1, SCN_SRC_MATERIAL = 2 };
typedef struct scn_source {
    int32_t type;            /* SCN_SRC_* */
    int32_t isotropic;       /* POINT: 1 = isotropic, 0 = fixed `dir` */
    int32_t G;               /* group when n_probG == 0 (MATERIAL: always G) */
    int32_t mat;             /* MATERIAL: source material index (1-based) */
    int32_t n_probG;         /* POINT: length of probG (= number of groups) or 0 */
    int32_t pad;
    double r[3];             /* POINT position */
    double dir[3];           /* POINT direction */
    double bottom[3];        /* MATERIAL bounding box */
    double top[3];
    double t_low, t_high;    /* MATERIAL boundingTime */
    const double* probG;     /* POINT: group distribution or NULL */
} scn_source;

/* One fixed-source solve: fixedSourcePhysicsPackage%run and %cycles
 * (PhysicsPackages/fixedSourcePhysicsPackage_class.f90:135-289).  Batch b (0-based) samples
 * source particle j (1-based) from the seed strided by 2*pop*b + j and runs history j on the seed
 * strided by 2*pop*b + pop + j, exactly the strides the package applies.                        */
typedef struct scn_fixed_request {
    const scn_problem* problem; /* its `transport` must match the batch engine's */
    const scn_source* source;
    int32_t pop;             /* histories per batch ('pop') */
    int32_t n_batches;       /* maximum number of batches ('cycles') */
    int64_t seed;
    int64_t stop_bin;        /* bin whose relative standard deviation may end the solve early; -1: never */
    double stop_rel;         /* target relative standard deviation for stop_bin (after >= 2 batches) */
    int32_t record;          /* 1: keep each history's final RNG state and collision count */
    int32_t buffer;          /* 'buffer': depth of a history's LIFO buffer of fission secondaries (0: 50) */
    int32_t window;          /* batches of this solve transported together in one engine cycle (0: 1,
                                at most 16; more than 1 needs an engine with n_launch 1).  They are
                                closed in order, so results and the early-stop point are those of
                                one batch per cycle; batches past an early stop are discarded. */
    int32_t adapt;           /* 0: every cycle transports `window` batches.  1-100 (with a stop_bin): the
                                first cycle transports 2, each later one adapt% of the batches the stop
                                bin's trend still asks for, adapt/100 (N (rel / stop_rel)^2 - N) with N
                                batches closed (at least 1, at most `window`), so that few batches are
                                transported past the stop; results and the stop batch do not change.
                                (No counterpart in the reference.) */
} scn_fixed_request;

enum { SCN_SLOT_FREE = 0, SCN_SLOT_RUNNING = 1, SCN_SLOT_DONE = 2 };

typedef struct scn_fixed_result {
    int32_t state;           /* SCN_SLOT_* */
    int32_t n_batches;       /* batches run (scoreMemory batchN) */
    int64_t n_bins;
    int64_t n_histories;
    int64_t n_collisions;    /* real collisions */
    int64_t n_segments;      /* tracking events (ST segments / DT tentative flights) */
} scn_fixed_result;

typedef struct scn_batch_settings {
    int32_t transport;       /* SCN_TRANSPORT_ST or _DT, shared by every resident problem */
    int32_t max_slots;       /* problems resident at once */
    int64_t hbm_budget;      /* device bytes for problem tables, particle queues and score memory */
    int32_t tile;            /* histories per work tile (multiple of 64); 0 = default */
    int32_t k0;              /* (flight, collision) pairs per particle in the first launch; 0 = default */
    int32_t n_launch;        /* launches per cycle, the last runs to completion; 0 = default */
    int32_t use_graph;       /* 1: replay each cycle's launch sequence from a captured hipGraph */
    int32_t refill;          /* idle lanes a wave gathers before it starts new histories (1..64, 0 = default):
                                a start (source sampling, RNG stride, placement) runs once for many lanes
                                instead of once for every lane that frees up; results do not change */
    int32_t pad0;
} scn_batch_settings;

/* ---- counters of the engine (cumulative since creation or the last reset) ----------------- */
typedef struct scn_stats {
    double transport_kernel_ms; /* summed HIP-event time of the transport kernel launches */
    int64_t transport_launches;
    int64_t n_histories;
    int64_t n_segments;          /* ST surface-tracking segments / DT tentative flights */
    int64_t n_collisions;        /* real collisions */
    int64_t n_sites;             /* fission sites banked */
    int64_t n_scores;            /* collision-clerk score additions */
    int64_t n_cycles;
    double cycle_ms;             /* wall time of whole cycles (transport + cycle close) */
} scn_stats;

/* ---- entry points --------------------------------------------------------------------------- */

/* Build an engine from flat tables (replaces geometry/nuclear-data/tally construction in
 * eigenPhysicsPackage_class.f90:507-590).  capacity = largest bank the handle will transport. */
int scn_create(const scn_problem* problem, int32_t capacity, void** engine);
void scn_destroy(void* engine);
const char* scn_last_error(const void* engine);
int scn_device_info(char* name, int32_t name_len, int32_t* n_cu);

/* One eigenvalue cycle of transport: the body of eigenPhysicsPackage%cycles between
 * reportCycleStart and reportCycleEnd (eigenPhysicsPackage_class.f90:208-252).
 *   bank_in   : thisCycle%prisoners(1:n_in); history n uses rng_state strided by n.
 *   k_eff     : k_new loaded into every history (eigen...:227).
 *   bank_out  : nextCycle in broodID order (what normSize_Repr's sortByBroodID produces).
 *   bin_scores: this cycle's BIN values of the collision clerks (length n_bins), or NULL.
 *   keff_scores[4]: keffImplicitClerk IMP_PROD, SCATTER_PROD, IMP_ABS, ANA_LEAK for the cycle.
 * The RNG state is not advanced: the caller keeps doing the strides, as the package does.   */
int scn_eigen_cycle(void* engine, const scn_particle_state* bank_in, int32_t n_in, int64_t rng_state,
                    double k_eff, scn_particle_state* bank_out, int32_t bank_out_cap, int32_t* n_out,
                    double* bin_scores, double* keff_scores);

/* Population control of one cycle: particleDungeon%normSize_Repr (particleDungeon_class.f90:
 * 431-602) on a broodID-sorted bank, single rank.  Works in place on bank[0..*n) (capacity cap). */
int scn_norm_size_repr(void* engine, scn_particle_state* bank, int32_t* n, int32_t cap,
                       int32_t tot_pop, int64_t rng_state);

/* Whole power iteration with banks and tallies resident in HBM: eigenPhysicsPackage%run
 * (eigenPhysicsPackage_class.f90:135-159) from an already generated initial bank. */
int scn_eigen_run(void* engine, const scn_eigen_settings* settings, const scn_particle_state* initial_bank,
                  scn_eigen_result* result);

/* The same power iteration split into calls, so a caller can time or interleave cycles:
 * start = run() up to the first cycle, step = the next n cycles (inactive, then active; *done
 * receives how many ran), finish = collectResults.  scn_eigen_run = start + step(all) + finish. */
int scn_eigen_start(void* engine, const scn_eigen_settings* settings, const scn_particle_state* initial_bank);
int scn_eigen_step(void* engine, int32_t n_cycles, int32_t* done);
int scn_eigen_finish(void* engine, scn_eigen_result* result);

/* Initial fission source: fissionSource%sampleParticle for particles 1..n
 * (ParticleObjects/Source/fissionSource_class.f90:149-271, source_inter.f90:98-119);
 * bottom/top = sampling box, g_in = the group passed to fissionMG%sampleOut. */
int scn_fission_source(void* engine, int32_t n, int64_t rng_state, const double bottom[3],
                       const double top[3], int32_t g_in, int32_t attempts, scn_particle_state* bank);

/* Fixed-source engine with resident problem slots (batched and served solves).
 * A cycle advances every running slot by one batch (fixedSourcePhysicsPackage_class.f90:187-275):
 * all their histories share one launch sequence, each wave of the persistent transport kernel
 * taking tiles of one problem's queue.  Every slot keeps its own tables, RNG stream and score
 * memory, so a solve gives the same histories alone as in a mix, and tallies equal to summation
 * order.  Fission sites of a history are detained in its LIFO buffer and transported after it on
 * the same RNG stream (fixedSourcePhysicsPackage_class.f90:214-243); more than `buffer` pending
 * secondaries is an error, as in particleDungeon%detain.                                       */
int scn_batch_create(const scn_batch_settings* settings, void** batch);
void scn_batch_destroy(void* batch);
const char* scn_batch_last_error(const void* batch);
/* admit a solve: returns SCN_ERR_BUSY (no slot or not enough of the HBM budget free) or SCN_OK
 * with its slot id.  The solve starts with the next cycle. */
int scn_batch_add(void* batch, const scn_fixed_request* request, int32_t* slot);
/* run n_cycles cycles (or fewer if no slot is running); *n_running = slots still running */
int scn_batch_cycle(void* batch, int32_t n_cycles, int32_t* n_running);
/* state of a slot; for a DONE slot also the score results: bins [n_bins*2] (mean, sd) as
 * scoreMemory%getResult gives them, and the recorded histories (n_batches*pop, batch-major) when
 * the request asked for them.  Pointers may be NULL. */
int scn_batch_query(void* batch, int32_t slot, scn_fixed_result* result, double* bins, uint64_t* hist_rng,
                    int32_t* hist_coll);
/* SCN_SLOT_* of slots 0..n-1 in one call (n <= max_slots) */
int scn_batch_states(void* batch, int32_t* states, int32_t n);
/* free a DONE slot (its memory returns to the budget) */
int scn_batch_release(void* batch, int32_t slot);
/* batches of a running solve transported in each of the next cycles, 1 .. the window it was admitted
 * with.  The batches are closed in order as before, so the results and the early-stop batch do not
 * change; a serving loop narrows the window of a solve that is close to its stop target so that
 * batches past the stop are not transported.  (No counterpart in the reference.) */
int scn_batch_set_window(void* batch, int32_t slot, int32_t window);
/* HBM-budget bytes that the per-lane buffers of secondaries take once a solve with fission or weight
 * windows is resident (`buffer` entries per lane, 0 = the default 50; one block shared by every such
 * solve, sized by the deepest buffer); -1 without a device.  (No counterpart in the reference.) */
int64_t scn_batch_secondary_bytes(int32_t transport, int32_t buffer);
/* scoreMemory%getResult of one bin (mean, sd) over the batches closed so far, for a running or done
 * solve; *n_batches receives that count (result 0, 0 before the first batch closes) */
int scn_batch_bin(void* batch, int32_t slot, int64_t bin, double result[2], int32_t* n_batches);
/* device time and counters of the batch engine */
int scn_batch_stats(void* batch, scn_stats* out);

/* Population control that allocates nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (serving loops and harnesses that forbid hipMalloc, code that
 * must not touch the allocator).  It is scn_norm_size_repr on DEVICE banks with no engine handle:
 * bank_in[0..n) must already be in sortByBroodID order with broodIDs in [0, SCN_NOALLOC_MAX_SITES
 * + 1] (checked on the device; SCN_ERR_ARG otherwise), bank_out receives the tot_pop sites
 * (out_cap >= tot_pop), *n_out their number.  Its scratch (draws, flags, prefix sums, brood
 * bounds, sort storage) is one static device array of fixed size on each GPU the library runs
 * on, held until the process ends; n above SCN_NOALLOC_MAX_SITES is rejected with SCN_ERR_ARG.
 * Calls on one GPU share that array: calls that overlap in time on one GPU race and must be
 * serialised by the caller.  stream: a hipStream_t, or NULL for the default stream; the call
 * returns after the stream has finished the pass.  Errors: scn_last_error(NULL).              */
#define SCN_NOALLOC_MAX_SITES (1 << 20)
int scn_norm_size_repr_noalloc(const scn_particle_state* bank_in, int32_t n, scn_particle_state* bank_out,
                               int32_t out_cap, int32_t tot_pop, int64_t rng_state, void* stream, int32_t* n_out);

/* Engine counters and timings (no counterpart in the reference; SCONE only times transport). */
int scn_get_stats(void* engine, scn_stats* out);
int scn_reset_stats(void* engine);

#ifdef __cplusplus
}
#endif

#endif /* SCONE_GPU_H */
