/* gpucell/gpucell.h -- C ABI of libgpucell_hip.so, the MI355X (gfx950) SDPD
 * step library.
 *
 * Part 1 replaces, one for one, the 20 live host entry points of the
 * reference's kernel layer (namespaces cupar, src/ParticleGPU.cuh:8-37, and
 * cudiv, src/ConfigDomain.cuh:9-18).  Argument order, element types and the
 * meaning of every array are the reference's; pointers are device pointers
 * except the two optional host outputs (h_sdpdforce, h_defforce).  Each
 * function returns 0 on success and a nonzero gc status otherwise (the
 * reference aborts through CHECK, src/erro.cuh:6-20; the header-only wrapper
 * gpucell/compat.hpp restores that behaviour for drop-in host code).
 *
 * Part 2 is the fused engine: the same step (Particle::UpdatePar,
 * src/Particle.cu:1310-1329) as a handful of fused kernels over
 * library-owned state, with an FP64 position accumulator.
 *
 * All launches go to the stream set by gc_SetStream (default: stream 0).
 * Calls are not thread-safe with respect to each other.
 */
#ifndef GPUCELL_GPUCELL_H
#define GPUCELL_GPUCELL_H

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum {
    GC_OK = 0,
    GC_ERR_INVALID = -1,     /* bad argument (null pointer, negative size) */
    GC_ERR_NOCONSTS = -2,    /* gc_ConstsDataUp was never called */
    GC_ERR_GRID = -3,        /* cube grid unusable (periodic axis < 3 cubes,
                                cube edge < CutRadius, MaxCut > CutRadius) */
    GC_ERR_NOMEM = -4,       /* device allocation of library scratch failed */
    GC_ERR_TOPOLOGY = -5,    /* membrane ids out of range */
    GC_ERR_HIP = -6,         /* a HIP runtime call failed (see message) */
    GC_ERR_NONFINITE = -7    /* engine detected a non-finite state */
};

/* Last error message of this library on the calling host thread (never
 * NULL; valid until that thread's next call into the library). */
const char* gc_GetLastError(void);
/* Library version string. */
const char* gc_Version(void);
/* Stream for all subsequent launches (a hipStream_t; NULL = stream 0). */
int gc_SetStream(void* stream);
/* Wait for the current stream. */
int gc_Synchronize(void);

/* ---- Part 1: reference entry points ---------------------------------- */

/* cupar::ConstsDataUp, src/ParticleGPU.cuh:10; called at Particle.cu:662.
 * Keeps the parameter block for every later launch. */
int gc_ConstsDataUp(const Structconst* consts);

/* No reference counterpart: settings the reference hard-codes.  Optional;
 * gc_DefaultExtConsts fills the defaults used when it is never called. */
void gc_DefaultExtConsts(gc_ext_consts* ext);
int gc_SetExtConsts(const gc_ext_consts* ext);
int gc_GetExtConsts(gc_ext_consts* ext);

/* Host-side grid sizing, Particle::ConfigDomain (src/Particle.cu:548-571):
 * per axis nCube = floor(L_pad / CutRadius), OuterRegion = L_pad,
 * invWid = nCube / OuterRegion, where L_pad = L on periodic axes and
 * L + 2*CutRadius on walled ones.  Returns SizeCube through *sizecube. */
int gc_ConfigDomain(const float bbox[6], float cutradius, const int periodic[3],
                    int ncube[3], float outerregion[3], float invwid[3],
                    int* sizecube);

/* cudiv::PreConList, src/ConfigDomain.cuh:11; call Particle.cu:668.
 * Per slot p: cc[p] = cube coordinates, cubepart[p] = cube key
 * (x fastest), oldpos[sortpart[p]] = p.  cutradius is accepted for
 * signature compatibility; the wall padding uses the constant block's
 * CutRadius (the reference passes CutRadius truncated to int). */
int gc_PreConList(int n, int cutradius, tfloat3* d_totalcoor, float* d_bboxsph,
                  float* d_invwid, int* d_ncube, tint3* d_cc, int* d_cubepart,
                  int* d_sortpart, int* d_oldpos);

/* cudiv::Sort, src/ConfigDomain.cuh:12; call Particle.cu:669.
 * Sorts (keys, values) by key, ties by ascending value; keys >= 0.
 * Deterministic (counting sort); synchronises once to size its histogram. */
int gc_Sort(int n, int* keys, int* values);

/* gc_Sort without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same results as gc_Sort.  Its
 * scratch (key range words, key counts and offsets, ranks, scatter arrays,
 * scan workspace) is one fixed device array per GPU of about 84 MB, held
 * until the process ends; sizes are limited to n <= GC_NOALLOC_MAX_N
 * elements and keys < GC_NOALLOC_MAX_K, larger ones return GC_ERR_INVALID.
 * Every call on a GPU shares that array: calls that overlap on one GPU
 * (two host threads, or two streams) race.  Like gc_Sort it synchronises
 * the library stream twice (key range, population check). */
#define GC_NOALLOC_MAX_N (1 << 22)
#define GC_NOALLOC_MAX_K (1 << 22)
int gc_SortNoAlloc(int n, int* keys, int* values);

/* cudiv::CalBeginEndCube, src/ConfigDomain.cuh:13; call Particle.cu:672.
 * beginendcube[k] = [first slot, last slot + 1) of cube k in the sorted
 * key array; an empty cube has x == y (== the slot where it would start). */
int gc_CalBeginEndCube(int n, int sizecube, int* d_cubepart, tint2* d_beginendcube);

/* cudiv::SortBasicArrays, src/ConfigDomain.cuh:14; call Particle.cu:677.
 * new[s] = old[oldpos[sortpart[s]]] for coor, cc, velrhop, velhalf,
 * totalforce, press. */
int gc_SortBasicArrays(int n, int* d_sortpart, int* d_oldpos, tfloat3* d_totalcoor,
                       tint3* d_cc, tfloat4* d_velrhop, tfloat3* d_velhalf,
                       tfloat3* d_totalforce, float* d_press);

/* cudiv::SortArraysWithCells, src/ConfigDomain.cuh:15; call Particle.cu:675.
 * As gc_SortBasicArrays plus nthrobound, and curpos[sortpart[s]] = s. */
int gc_SortArraysWithCells(int n, int* d_sortpart, int* d_oldpos, int* d_curpos,
                           tfloat3* d_totalcoor, tint3* d_cc, tfloat4* d_velrhop,
                           tfloat3* d_velhalf, tfloat3* d_totalforce, float* d_press,
                           tint3* d_nthrobound);

/* cupar::UpdateFluidCoor, src/ParticleGPU.cuh:11; call Particle.cu:1333.
 * r += dt v + dt^2/2 F/m for ids >= nFluidStart. */
int gc_UpdateFluidCoor(int n, int* d_sortpart, tfloat3* d_totalforce,
                       tfloat3* d_totalcoor, tfloat4* d_velrhop);

/* cupar::AppFluidPeriodicBC, src/ParticleGPU.cuh:12; call Particle.cu:1338.
 * Wraps periodic axes into [lo, hi) and counts the crossings in nthrobound
 * (unwrapped = wrapped + nthrobound * L). */
int gc_AppFluidPeriodicBC(int n, int* d_sortpart, tfloat3* d_totalcoor,
                          tint3* d_nthrobound, float* d_bboxsdpd);

/* cupar::CalFluidDen, src/ParticleGPU.cuh:13; call Particle.cu:1343.
 * rho_i = m sum_j W(r_ij) over Bound2, fluid and membrane neighbours
 * (self included) for fluid and membrane ids; walls keep their density. */
int gc_CalFluidDen(int n, tfloat3* d_totalcoor, tfloat4* d_velrhop, tint3* d_cc,
                   int* d_ncube, float* d_outerregion, int* d_sortpart,
                   tint2* d_beginendcube);

/* cupar::PredictFluidVel, src/ParticleGPU.cuh:14; call Particle.cu:1348.
 * velhalf = v + dt/2 F/m, v = v + lambda dt F/m for ids >= nFluidStart. */
int gc_PredictFluidVel(int n, int* d_sortpart, tfloat4* d_velrhop,
                       tfloat3* d_velhalf, tfloat3* d_totalforce);

/* cupar::CalBound2Vel, src/ParticleGPU.cuh:15; calls Particle.cu:1317,1327.
 * Bound2 velocity = -(sum_j W_ij v_j)/(sum_j W_ij) over moving neighbours
 * (0 when there are none): a no-slip image of the adjacent flow. */
int gc_CalBound2Vel(int n, tfloat3* d_totalcoor, tfloat4* d_velrhop, tint3* d_cc,
                    int* d_ncube, float* d_outerregion, int* d_sortpart,
                    tint2* d_beginendcube);

/* cupar::EOS, src/ParticleGPU.cuh:16; call Particle.cu:1358.
 * P = c^2 rho, c = SoundSpeed1 (ids < nCellStart) or SoundSpeed2. */
int gc_EOS(int n, int* d_sortpart, tfloat4* d_velrhop, float* d_press);

/* cupar::CalFluidSDPDForce, src/ParticleGPU.cuh:17; call Particle.cu:1363.
 * Overwrites totalforce with the SDPD pair force (conservative, dissipative,
 * random) for ids >= nFluidStart, zero for walls.  Uses and advances the
 * noise counter.  h_sdpdforce: optional host copy (slot order), may be NULL. */
int gc_CalFluidSDPDForce(int n, tfloat3* h_sdpdforce, tfloat3* d_totalcoor,
                         tfloat4* d_velrhop, tint3* d_cc, int* d_ncube,
                         int* d_sortpart, float* d_outerreigon,
                         tint2* d_beginendcube, tfloat3* d_totalforce,
                         float* d_press);

/* cupar::CalFluidRepForce, src/ParticleGPU.cuh:18; call Particle.cu:1368.
 * Adds the Bound1 wall repulsion to ids >= nFluidStart. */
int gc_CalFluidRepForce(int n, tfloat3* d_totalcoor, tint3* d_cc, int* d_ncube,
                        int* d_sortpart, float* d_outerreigon,
                        tint2* d_beginendcube, tfloat3* d_totalforce);

/* cupar::SetFluidExtForce, src/ParticleGPU.cuh:19; call Particle.cu:1373.
 * Adds the body force (ConstExtForce along x, pattern from the ext block)
 * to ids >= nFluidStart when driveon != 0. */
int gc_SetFluidExtForce(int driveon, int n, int* d_sortpart, tfloat3* d_totalforce);
/* No reference counterpart: gc_SetFluidExtForce with the positions the
 * reverse-Poiseuille pattern needs (the reference entry has none). */
int gc_SetFluidExtForcePos(int driveon, int n, int* d_sortpart, tfloat3* d_totalcoor,
                           tfloat3* d_totalforce);

/* cupar::CorrectFliudVel, src/ParticleGPU.cuh:21; call Particle.cu:1379.
 * v = velhalf + dt/2 F/m for ids >= nFluidStart (Groot-Warren corrector);
 * the neighbour arguments are accepted and unused. */
int gc_CorrectFliudVel(int n, int* d_sortpart, tfloat3* d_totalforce,
                       tfloat3* d_velhalf, tfloat4* d_velrhop, tfloat3* d_totalcoor,
                       tint3* d_cc, int* d_ncube, float* d_outerreigon,
                       tint2* d_beginendcube);

/* cupar::CalGeoPara, src/ParticleGPU.cuh:25-26; calls Particle.cu:476,1322.
 * Per triangle normal/area/centre from unwrapped vertices; per cell area
 * and volume reduced in FP64; per-vertex area (1/3 of incident triangles).
 * stepcount == -1 also stores the reference state. */
int gc_CalGeoPara(long stepcount, int ntri, int ncell, float* d_bbox, int* d_curpos,
                  tint3* d_tri, tint2* d_beginendtri, tint3* d_nthrobound,
                  tfloat3* d_totalcoor, tfloat3* d_trinordir, float* d_triarea,
                  float* d_reftriarea, tfloat3* d_tricenter, float* d_refarea_para,
                  float* d_refvol_para, float* d_area_para, float* d_vol_para,
                  float* d_area, float* d_vol, float* d_pararea);

/* cupar::SetMechPara, src/ParticleGPU.cuh:28-29; call Particle.cu:479.
 * Per edge: rest length, WLC maximum length, persistence length, POW
 * coefficient and exponent from the shear modulus; bending coefficient
 * 2 k_c / sqrt(3); rest dihedral angle from the initial geometry. */
int gc_SetMechPara(int ncell, int nedge, int6* d_edge, int* d_curpos,
                   float* d_refedgelen, float* d_maxrefedgelen, float* d_bendmodpara,
                   float* d_shearmodpara, float* d_hbending, float* d_powerindex,
                   float* d_perlenWLC, float* d_hpow, float* d_reftriangle,
                   tfloat3* d_trinordir, tfloat3* d_tricente, tfloat3* d_coor);

/* cupar::CalCellDefForce, src/ParticleGPU.cuh:30-33; call Particle.cu:1385.
 * WLC-POW springs, dihedral bending, local/global area and volume
 * penalties; d_defforce[v] = membrane force of vertex v (id order),
 * totalforce[curpos[id]] += it.  Accumulation is per vertex in a fixed
 * order (no float atomics).  h_defforce: optional host copy, may be NULL. */
int gc_CalCellDefForce(int ncell, unsigned nedge, unsigned ntri, unsigned ntotalpar,
                       unsigned ncellpar, float* d_bbox, int6* d_edge, tint3* d_tri,
                       tfloat3* d_coor, tfloat3* h_defforce, tfloat3* d_defforce,
                       tint3* d_nthrobound, float* d_maxrefedgelen,
                       float* d_powerindex, float* d_perlenWLC, float* d_hpow,
                       tfloat3* d_trinordir, tfloat3* d_tricenter,
                       float* d_reftriangle, float* d_hbending, float* d_triarea,
                       float* d_reftriarea, float* d_hlarea_para,
                       float* d_hgarea_para, float* d_area_para,
                       float* d_refarea_para, float* d_hvol_para, float* d_vol_para,
                       float* d_refvol_para, tfloat3* d_totalforce, int* d_sortpart,
                       int* d_curpos);

/* cupar::CalCellAggForce, src/ParticleGPU.cuh:34-35; call Particle.cu:1390.
 * Morse force between membrane vertices of different cells within MaxCut;
 * d_aggforce[v] (id order) = it, totalforce at the vertex's slot += it. */
int gc_CalCellAggForce(unsigned ncell, unsigned n, unsigned ncellpar,
                       tfloat3* d_totalcoor, tfloat3* d_aggforce, tint3* d_cc,
                       int* d_ncube, int* d_sortpart, int* d_celltab,
                       float* d_outerreigon, tint2* d_beginendcube,
                       tfloat3* d_totalforce, float* d_pararea);

/* The membrane kernels derive a vertex -> (edge, triangle) incidence table
 * from d_edge/d_tri on first use and cache it per (pointer, count).  Call
 * this after changing topology in place.  No reference counterpart. */
int gc_InvalidateTopology(void);

/* ---- Part 2: fused engine -------------------------------------------- */

/* A whole case in load (id) order, as the reference's loaders leave it
 * (Particle::LoadParInfo, src/Particle.cu:29-427).  Host pointers. */
typedef struct {
    int nBound1Par, nBound2Par, nFluidPar;   /* Particle.cuh:13 */
    int nCell, nAllTri, nAllEdge;            /* Particle.cuh:24 */
    int nCellPar;                            /* TypesDef.cuh:56 */
    const tfloat3* coor;          /* [nTotalPar] positions */
    const tfloat3* vel;           /* [nTotalPar] initial velocity or NULL (= 0) */
    const tint3* tri;             /* [nAllTri] 1-based membrane vertex ids */
    const int6* edge;             /* [nAllEdge] 1-based vertex and triangle ids */
    const int* celltab;           /* [nCellPar] cell of each membrane vertex */
    const tint2* beginendtri;     /* [nCell] inclusive triangle range */
    const float* shearmod_para;   /* [nAllEdge] mu0 */
    const float* bendingmod_para; /* [nAllEdge] k_c */
    const float* hlarea_para;     /* [nAllTri] k_d */
    const float* hgarea_para;     /* [nAllTri] k_a */
    const float* hvol_para;       /* [nAllTri] k_v */
    float bbox[6];                /* BBox (main.cuh:12) */
    Structconst consts;           /* counts inside are recomputed */
    gc_ext_consts ext;
    /* Verlet skin of the engine's neighbour lists (no reference
     * counterpart).  0: the reference's per-step scheme (re-sort and cube
     * scan every step; cube side CutRadius).  > 0: cube side CutRadius +
     * skin, lists of every pair within CutRadius + skin rebuilt (with the
     * re-sort) only when some pair may have moved apart by more than skin
     * since the last rebuild (a displacement beyond skin / 2, and a
     * displacement bounding box wider than skin), decided on the device per
     * replica; the density and force passes iterate the lists.  Every
     * replica of an ensemble must use the same mode (skin 0 for all, or > 0
     * for all). */
    float skin;
} gc_case_desc;

typedef struct gc_engine gc_engine;

/* Allocates, uploads and initialises (Particle::SetPar, src/Particle.cu:
 * 471-483: zero velocity unless given, rho = DenInit, first neighbour build,
 * reference geometry, mechanical parameters).  NULL on error (*status set). */
gc_engine* gc_EngineCreate(const gc_case_desc* desc, int* status);
void gc_EngineDestroy(gc_engine* e);
/* nsteps x Particle::UpdatePar with DriveOn = driveon (Particle.cu:1302-1308). */
int gc_EngineStep(gc_engine* e, int nsteps, int driveon);
/* Same, captured once into a HIP graph and replayed (no host work per step). */
int gc_EngineStepGraph(gc_engine* e, int nsteps, int driveon);

/* Fault bits an engine's kernels raise (sticky until gc_EngineSetState or
 * gc_EngineLoad): a non-finite force, or a membrane edge stretched to or
 * past its WLC maximum length l_max (x = l / l_max >= 1, where the spring
 * force is singular).  Export and checkpoint calls return GC_ERR_NONFINITE
 * while any bit is set (the data are still copied out for inspection). */
enum { GC_FAULT_NONFINITE = 1, GC_FAULT_WLC = 2 };

typedef struct {
    long step;                    /* steps taken */
    uint64_t noise_counter;
    int nTotalPar, nCellPar, nCell, sizecube;
    int ncube[3];
    float outerregion[3], invwid[3];
    int fault;                    /* GC_FAULT_* bits raised so far */
    int nl_overflow;              /* particles of the last step (skin mode:
                                     of the last list build) with more
                                     neighbours than the list holds (their
                                     force and Morse passes scan the cubes) */
    int nrebuild;                 /* skin mode: list rebuilds so far, summed
                                     over the replicas (0 in per-step mode) */
    float skin;                   /* the engine's Verlet skin (0: per step) */
} gc_engine_info;
/* Synchronises; returns GC_OK whatever the fault bits (they are in info). */
int gc_EngineInfo(const gc_engine* e, gc_engine_info* info);

/* Consistency of the slot maps of an engine's particle order (a debugging
 * aid; no counterpart in the reference, which keeps CurPos for every
 * particle): sp[s] = id at slot s must be a permutation of [0, n);
 * curpos[id] = slot of membrane vertex id (ids in [nc, n)) with
 * sp[curpos[id]] == id; act = the nmov slots the force pass visits: the
 * slots of the moving particles (ids >= nf) in increasing order, or, with
 * moving walls (nmov == n), every slot in order.  Host code only (no GPU).
 * GC_OK, or GC_ERR_INVALID with a message naming the first violation. */
int gc_CheckSlotMaps(int n, int nf, int nc, int nmov, const int* sp, const int* curpos, const int* act);
/* The engine's current maps (valid after a step) copied to the host and
 * checked as above.  Setting GPUCELL_CHECK_SLOTS=1 before gc_EngineCreate
 * runs the same check after the re-sort of every eager step. */
int gc_EngineCheckSlotMaps(gc_engine* e);

/* State copied to host in id order; any pointer may be NULL. */
typedef struct {
    double* pos;        /* [N*3] FP64 position accumulator (wrapped) */
    tfloat3* coor;      /* [N] FP32 positions the kernels read */
    tfloat4* velrhop;   /* [N] */
    tfloat3* velhalf;   /* [N] */
    tfloat3* force;     /* [N] total force of the last step */
    float* press;       /* [N] */
    tint3* nthrobound;  /* [N] */
    float* area;        /* [nCell] */
    float* vol;         /* [nCell] */
    tfloat3* defforce;  /* [nCellPar] */
    tfloat3* aggforce;  /* [nCellPar] */
} gc_host_state;
int gc_EngineGetState(gc_engine* e, gc_host_state* out);
/* Overwrites the dynamic state (id order) and re-sorts; NULL keeps a field. */
int gc_EngineSetState(gc_engine* e, const gc_host_state* in, long step,
                      uint64_t noise_counter);

/* Slot-order neighbour structure of the last build (device -> host). */
typedef struct {
    int* sortpart;      /* [N] slot -> id */
    int* cubepart;      /* [N] cube key per slot */
    tint3* cc;          /* [N] */
    int* curpos;        /* [N] id -> slot */
    tint2* beginendcube;/* [sizecube] */
} gc_host_cells;
int gc_EngineGetCells(gc_engine* e, gc_host_cells* out);

/* Triangle normals and centres of the last step (nAllTri each). */
int gc_EngineGetTriangles(gc_engine* e, tfloat3* trinordir, tfloat3* tricenter);
/* nBound2Start and nFluidStart of the engine's parameter block. */
int gc_EngineRanges(const gc_engine* e, int* nbound1, int* nbound12);

/* ---- case decks and Tecplot output (host only) ------------------------ */

/* Reads the reference's four input decks from `dir` (SDPDParCoor.txt,
 * SDPDPhyPara.txt, CELLParStruct.txt, CELLPhyPara.txt; formats of
 * Particle::LoadFBPar/LoadSDPDPara/LoadCellStruct/LoadCellPara,
 * src/Particle.cu:45-427), with the compile-time constants of
 * src/main.cuh:6-16.  Arrays are allocated; release with gc_FreeCase. */
int gc_ReadCase(const char* dir, gc_case_desc* out);
void gc_FreeCase(gc_case_desc* d);
/* Writes a case as the four decks (fields narrower than 20 columns). */
int gc_WriteCase(const char* dir, const gc_case_desc* d);
/* One frame of Particle::OutputPar (src/Particle.cu:681-1300) under
 * dir/Result/: FluidCoor, FluidField, CellCoor, CellForce (+TF_Cell<i>),
 * CellGeo (Geo_Cell<i>), CellNorDir; stepcount == 0 starts the per-cell
 * time series, later frames append. */
int gc_EngineWriteTecplot(gc_engine* e, const char* dir, long stepdata, long stepcount,
                          float timestep, long stepoutput, const tint3* tri,
                          const tint2* beginendtri, const int* celltab);

/* Checkpoint / restart (no reference counterpart: the reference's outputs
 * keep 5 digits).  Save writes the dynamic state (FP64 positions, Velrhop,
 * VelHalf, TotalForce, nThroBound, step, noise counter) in id order; Load
 * restores it into an engine created from the same case, after which the
 * run continues bit for bit. */
int gc_EngineSave(gc_engine* e, const char* path);
int gc_EngineLoad(gc_engine* e, const char* path);

/* ---- autotuner (no reference counterpart) -----------------------------
 * Launch shapes of the hot kernels: integrate (block 64/128/256), density
 * (block 64/128/256 x LDS row stage 0/128/192/320 particles per wave), SDPD
 * force (block 64/128/256), membrane force (group 64/128/256).  Every shape
 * computes bit-identical results; only the time differs.  Autotune times
 * each variant over `steps` eager steps with HIP events on the engine
 * stream, keeps the fastest per family, and restores the state it started
 * from (the trajectory is unchanged).  One report row per variant. */
typedef struct {
    int n;
    char family[64][24];
    int variant[64];
    char shape[64][32];
    float us[64];
    int chosen[64];
} gc_tune_report;
int gc_EngineAutotune(gc_engine* e, int steps, int driveon, gc_tune_report* out);
/* Current variant per family (4 ints: integrate, density, force, membrane). */
int gc_EngineGetTuning(const gc_engine* e, int* variants);
int gc_EngineSetTuning(gc_engine* e, const int* variants);

/* ---- batched ensembles (no reference counterpart) ---------------------
 * One engine over nrep independent suspensions of any sizes and physics
 * (parameter sweeps, ensemble statistics): the same kernels and graph step
 * all of them, so small suspensions share the launches instead of leaving
 * the GPU idle.  Every replica evolves bit for bit as gc_EngineCreate on its
 * own descriptor (its own ids and seed drive its pair noise); all replicas
 * must start from the same noise counter.  gc_EngineStep/StepGraph/Profile/
 * Info/Destroy apply to the whole batch; state is read per replica.
 * gc_EngineSetState/Save/Load/WriteTecplot refuse a batch of more than one. */
gc_engine* gc_EnsembleCreate(const gc_case_desc* descs, int nrep, int* status);
/* Replica `rep`'s state in its own id order (arrays sized for that replica). */
int gc_EnsembleGetState(gc_engine* e, int rep, gc_host_state* out);
/* Number of replicas (1 for gc_EngineCreate; 0 for NULL). */
int gc_EnsembleCount(const gc_engine* e);

/* Per-kernel averages of the last gc_EngineProfile call, in microseconds. */
typedef struct {
    int nkernels;
    char names[16][32];
    float us[16];
} gc_engine_profile;
/* Times nsteps steps kernel by kernel with HIP events on the engine stream. */
int gc_EngineProfile(gc_engine* e, int nsteps, int driveon, gc_engine_profile* out);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* GPUCELL_GPUCELL_H */
