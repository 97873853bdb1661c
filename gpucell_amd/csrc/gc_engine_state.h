// gc_engine_state.h -- the engine's library-owned state, shared by the step
// (gc_engine.hip) and the membrane time stepping (gc_substep.hip).
#pragma once

#include <vector>

#include "gc_kernels.h"

namespace gc {

template <class T>
int dalloc(T** p, size_t count)
{
    if (hipMalloc((void**)p, sizeof(T) * (count ? count : 1)) != hipSuccess) {
        *p = nullptr;
        return set_error(GC_ERR_NOMEM, "engine allocation failed");
    }
    return GC_OK;
}

template <class T>
int upload(T* d, const T* h, size_t count, hipStream_t s)
{
    if (!count) return GC_OK;
    return hip_check(hipMemcpyAsync(d, h, sizeof(T) * count, hipMemcpyHostToDevice, s), "engine H2D");
}

// Per-replica bookkeeping on the host.
struct RepMeta {
    int nb1, nb2, nfl, nm, ncell, ntri, nedge;   // the replica's own counts
    int vtx_off, cell_off, tri_off, edge_off;    // its first membrane vertex / cell / triangle / edge
    Structconst consts;                          // its own id ranges
};

// Launch shapes picked by the autotuner (gc_EngineAutotune): one variant
// index per kernel family; every variant computes bit-identical results.
enum { TUNE_INTEGRATE = 0, TUNE_DENSITY = 1, TUNE_FORCE = 2, TUNE_MEMBRANE = 3, TUNE_N = 4 };
struct TuneCfg {
    // integrate 256, density 256 x 192, force 256, membrane per cell (groups of
    // 64 threads in three passes when a cell does not fit)
    int v[TUNE_N] = {2, 10, 2, 3};
};
constexpr int TUNE_NVARIANT[TUNE_N] = {3, 12, 3, 4};

// Skin mode: displacement statistics per workgroup since the last list
// rebuild (the largest |D|^2, then per axis the largest max(D, 0) and
// max(-D, 0)), written by the advance pass and the membrane substeps.
constexpr int DSTAT_N = 7;

// Membrane substepping / surrogate state (gc_substep.hip); NULL while the
// membrane moves with the fluid step.
struct MembraneSub;
void ms_destroy(MembraneSub* m);

}  // namespace gc

struct gc_engine {
    int nrep = 1;
    std::vector<gc::GcRep> hreps;
    std::vector<gc::RepMeta> meta;
    int N = 0, nm = 0, ncell = 0, ntri = 0, nedge = 0, K = 0;
    Structconst consts;        // global type thresholds (physics of replica 0)
    gc_ext_consts ext;
    long step = 0;
    hipStream_t s = nullptr;
    int cur = 0;
    std::vector<void*> allocs;

    gc::GcRep* d_reps = nullptr; int* d_rep_of = nullptr; int* d_cell_rep = nullptr;
    float* d_bbox = nullptr;   // box of replica 0 (one suspension: the box)
    double* d_pos64[2]; tint3* d_nt[2]; int* d_sp[2];
    float4 *d_posid, *d_velrho, *d_vout;
    tfloat3 *d_coor, *d_velhalf, *d_force;
    int *d_key0, *d_key, *d_rank, *d_cnt, *d_offs, *d_tk, *d_tv, *d_to, *d_curpos;
    // moving particles: walls do not move, so the number of walls per cube is
    // fixed; wincl = its inclusive prefix over cubes; act[i] = slot of the
    // i-th moving particle in slot order (written by the gather), nmov of them
    int *d_wincl = nullptr, *d_act = nullptr;
    int nmov = 0;
    int moving_walls = 0;   // some suspension's Bound2 walls translate (Couette shear)
    int* d_tsum = nullptr;   // cube counts per scan tile
    int ntile = 0;
    int *d_nl, *d_nlc;   // neighbour lists of moving particles (density -> force)
    int nlcap = 64;      // list entries per particle (skin mode: 128)
    // Verlet skin mode (skin > 0): slots, keys and lists change only at a
    // rebuild; buffer 0 is always current (the rebuild gathers into buffer
    // 1 and copies back).  Per replica: the largest squared displacement
    // since the last rebuild from workgroups spanning replicas (float bits),
    // the rebuild flag of this step and the rebuild count; rbany = some
    // replica rebuilds this step (gates the rebuild kernels).
    float skin = 0.0f;
    unsigned* d_disp2 = nullptr;
    tfloat3* d_rbpos = nullptr;   // per slot: FP32 position at the last rebuild
    float* d_bmax = nullptr;      // per workgroup: largest squared displacement since the rebuild
    int* d_brep = nullptr;        // per workgroup: its replica (-1: several)
    int *d_rbflag = nullptr, *d_rbany = nullptr, *d_nrebuild = nullptr;
    int* d_ones = nullptr;   // an all-ones flag array (host-driven rebuilds of every replica)
    unsigned char* d_woff = nullptr;   // per slot: periodic wraps since the last rebuild (image code)
    // a batch: first slot and first cube of each replica (nrep + 1 entries,
    // the last = N or K); a replica's slot range is fixed (its particle
    // count is), so a rebuild re-sorts the rebuilding replicas only
    int *d_slotbase = nullptr, *d_cubebase = nullptr;
    int2* d_chunkrep = nullptr;          // per 64-slot chunk, then 64-cube chunk: first, last replica
    unsigned char* d_gact = nullptr;     // per chunk, set each rebuild: 1 if a rebuilding replica owns part
    bool skin_mode() const { return skin > 0.0f; }
    unsigned long long* d_ctr;
    int* d_fault = nullptr;    // GC_FAULT_* bits raised by the kernels (sticky)
    int* d_nlover = nullptr;   // density-pass list overflows of the last step
    // membrane
    tint3* d_tri; int6* d_edge; int* d_celltab; tint2* d_betri;
    float *d_mu0, *d_kc, *d_kd, *d_ka, *d_kv;
    float *d_l0, *d_lmax, *d_kb, *d_mexp, *d_p, *d_kp, *d_th0;
    float *d_cth0, *d_sth0;   // cos and sin of th0
    tfloat3 *d_trinor, *d_tricen; float *d_triarea, *d_reftriarea;
    float *d_refarea_para, *d_refvol_para, *d_area_para, *d_vol_para, *d_area, *d_vol, *d_pararea;
    tfloat3 *d_defforce, *d_aggforce;
    int4* d_cellr = nullptr;   // per-cell vertex and edge ranges (fused membrane kernel)
    bool cell_fused = false;   // every cell fits the fused kernel
    tfloat3* d_ef = nullptr;   // per-edge forces on the edge's four vertices
    tfloat3* d_tf = nullptr;   // per-triangle forces on the three corners
    double* d_mpos;
    int *d_es, *d_ee, *d_ts, *d_te;
    hipGraphExec_t gexec = nullptr;
    int g_driveon = -1;
    bool walls_dirty = false;
    bool check_slots = false;   // GPUCELL_CHECK_SLOTS=1: slot maps checked on the host after each eager re-sort
    bool any_morse = false, any_walls = false, any_bound1 = false;
    gc::MembraneSub* ms = nullptr;   // membrane substepping (mode 1/2)
    gc::TuneCfg tune;

    template <class T> int alloc(T** p, size_t n)
    {
        int st = gc::dalloc(p, n);
        if (!st) allocs.push_back((void*)*p);
        return st;
    }
    bool ens() const { return nrep > 1; }
};

namespace gc {
// One outer step of the membrane under substepping or the surrogate, after
// the fluid forces of the step (gc_substep.hip).
// membrane force of the engine's membrane (edge, triangle, vertex passes)
// with the arguments of k_def_force; P is the first replica's block
int membrane_force(gc_engine* e, const GcParams& P, const DefArgs& a, hipStream_t s);
// the fused per-cell kernel applies (tuning and cell sizes)
inline bool use_cell_kernel(const gc_engine* e) { return e->cell_fused && e->tune.v[TUNE_MEMBRANE] == 3; }
// geometry outputs of the engine for k_cell_membrane
CellArgs cell_args(gc_engine* e, long step, int round_geo);
int ms_step(gc_engine* e);
// Re-reads the membrane state from the engine arrays (after SetState).
int ms_reset(gc_engine* e);
// the membrane substep reads nothing back on the host (graph capture is safe)
bool ms_graph_ok(const gc_engine* e);
}  // namespace gc

struct gc_mlp;
namespace gc {
// Inference of the substepper network on stream s (gc_mlp.hip).
int mlp_forward(gc_mlp* m, const float* x, float* y, int batch, hipStream_t s);
}  // namespace gc
