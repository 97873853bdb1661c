// gc_kernels.h -- kernels shared between the entry-point translation units
// and the fused engine (defined in gc_sdpd.hip, gc_membrane.hip, gc_abi.hip).
#pragma once

#include <vector>

#include "gc_internal.h"

namespace gc {

// One suspension of a batched run (and the single suspension of an engine).
// Replicas are concatenated with the global id layout
//   [Bound1 of all | Bound2 of all | fluid of all | membrane of all],
// so P's type thresholds (nb2, nf, nc, N) are the global ones; the physics
// fields of P are the replica's own.  Each replica owns the global cube keys
// [cube_off, cube_off + nc0 nc1 nc2).
struct GcRep {
    GcParams P;
    int nc[3];          // cube grid
    float L[3];         // OuterRegion (period of periodic axes)
    float bbox[6];
    float invw[3];
    int cube_off;
    // cube order: x fastest within a (y, z) row; the rows are grouped into
    // tiles of 2^tys consecutive y (z, then y inside a tile; tiles in y), so
    // that the rows above and below a cube lie ~2^tys rows away in slot
    // order instead of ny (2^tys >= ny: the plain x, y, z order)
    int tys;
    int id_off[4];      // first global id of the replica's Bound1, Bound2, fluid, membrane
    int loc_start[4];   // the same blocks' first ids in the replica's own numbering
    // Verlet skin mode (skin > 0): neighbour lists hold every pair within
    // h + skin at the last rebuild (list_r2 = (h + skin)^2); a rebuild is
    // due once some displacement since the rebuild exceeds half_skin (skin
    // / 2 less a rounding margin) and the displacements' bounding box is
    // wider than twice that.  0 in per-step mode.
    float list_r2;
    float half_skin;
};

// Cube order: rows of nx cubes grouped into tiles of 2^tys y-rows holding
// at most CUBE_TILE_CUBES cubes, so that a cube's neighbours in the rows at
// z - 1 and z + 1 are at most that many cubes away in slot order instead of
// nx ny, and are still in the L2 of the XCD that sweeps them when they are
// read again (measured on the 10M-particle duct, 242 x 92 x 92 cubes, skin
// mode: 2.65 ms/step untiled, 2.70 / 2.63 / 2.60 / 2.57 / 2.59 / 2.62 with
// tiles of 2, 4, 8, 16, 32, 64 rows).  A grid whose whole y range fits one
// tile keeps the plain order.
constexpr int CUBE_TILE_CUBES = 4096;
__host__ __device__ inline int cube_tile_shift(int nx, int ny)
{
    int s = 0;
    while (s < 20 && (2 << s) * nx <= CUBE_TILE_CUBES) ++s;   // 2^s rows of nx cubes <= CUBE_TILE_CUBES
    if ((1 << s) >= ny) {   // one tile: 2^s >= ny
        s = 0;
        while ((1 << s) < ny) ++s;
    }
    return s;
}

// Index of row (y, z) of a replica's cubes (the key is cube_off + x +
// nx * row): tile t0 = the tile's first y, th its height (the last tile
// may be shorter).
__host__ __device__ inline int cube_row(const GcRep& R, int y, int z)
{
    const int t0 = (y >> R.tys) << R.tys;
    const int th = (R.nc[1] - t0) < (1 << R.tys) ? (R.nc[1] - t0) : (1 << R.tys);
    return t0 * R.nc[2] + z * th + (y - t0);
}

// Cube coordinates of key k (global key, cube_off included).
__host__ __device__ inline void cube_coords(const GcRep& R, int k, int& cx, int& cy, int& cz)
{
    k -= R.cube_off;
    cx = k % R.nc[0];
    const int r = k / R.nc[0];
    const int t0 = (r / (R.nc[2] << R.tys)) << R.tys;
    const int rem = r - t0 * R.nc[2];
    const int th = (R.nc[1] - t0) < (1 << R.tys) ? (R.nc[1] - t0) : (1 << R.tys);
    cz = rem / th;
    cy = t0 + (rem - cz * th);
}

// the replica's own id of global id `id` (used for the pair noise, so a
// replica draws the same numbers as when it runs alone).  Selects on
// constant indices: a runtime index into the struct would put it in scratch.
// The offsets come in by value: a conditional over lvalues selects their
// addresses, and inside a by-reference lambda capture such a pointer select
// keeps the captured locals in scratch.
__device__ __forceinline__ int local_id_of(int id, int nb2, int nf, int nc, int d0, int d1, int d2,
                                           int d3)
{
    return id - (id < nb2 ? d0 : (id < nf ? d1 : (id < nc ? d2 : d3)));
}

__device__ __forceinline__ int local_id(const GcRep& R, int id)
{
    return local_id_of(id, R.P.nb2, R.P.nf, R.P.nc, R.id_off[0] - R.loc_start[0],
                       R.id_off[1] - R.loc_start[1], R.id_off[2] - R.loc_start[2],
                       R.id_off[3] - R.loc_start[3]);
}

// gc_sdpd.hip
__global__ void k_density(int n, GcParams P, const tfloat3* __restrict__ coor,
                          tfloat4* __restrict__ velrhop, const tint3* __restrict__ cc,
                          const int* __restrict__ ncube, const float* __restrict__ outer,
                          const int* __restrict__ sortpart, const tint2* __restrict__ be);
__global__ void k_bound2_vel(int n, GcParams P, const tfloat3* __restrict__ coor,
                             tfloat4* __restrict__ velrhop, const tint3* __restrict__ cc,
                             const int* __restrict__ ncube, const float* __restrict__ outer,
                             const int* __restrict__ sortpart, const tint2* __restrict__ be);
__global__ void k_eos(int n, GcParams P, const int* __restrict__ sortpart,
                      const tfloat4* __restrict__ velrhop, float* __restrict__ press);
__global__ void k_sdpd_force(int n, GcParams P, const tfloat3* __restrict__ coor,
                             const tfloat4* __restrict__ velrhop, const tint3* __restrict__ cc,
                             const int* __restrict__ ncube, const int* __restrict__ sortpart,
                             const float* __restrict__ outer, const tint2* __restrict__ be,
                             tfloat3* __restrict__ force, const float* __restrict__ press,
                             const unsigned long long* __restrict__ ctr);
__global__ void k_rep_force(int n, GcParams P, const tfloat3* __restrict__ coor,
                            const tint3* __restrict__ cc, const int* __restrict__ ncube,
                            const int* __restrict__ sortpart, const float* __restrict__ outer,
                            const tint2* __restrict__ be, tfloat3* __restrict__ force);
__global__ void k_ext_force(int n, GcParams P, const int* __restrict__ sortpart,
                            const tfloat3* __restrict__ coor, tfloat3* __restrict__ force);
__global__ void k_correct(int n, GcParams P, const int* __restrict__ sortpart,
                          const tfloat3* __restrict__ force, const tfloat3* __restrict__ velhalf,
                          tfloat4* __restrict__ velrhop);

// gc_membrane.hip
struct DefArgs {
    const int6* edge; const tint3* tri; const tfloat3* coor; const tint3* nthro; const int* curpos;
    const float* lmax; const float* mexp; const float* p; const float* kp; const float* th0;
    const float* kb; const float* reftriarea; const float* kd; const float* ka; const float* Ac;
    const float* A0c; const float* kv; const float* Vc; const float* V0c;
    const int* es; const int* ee; const int* ts; const int* te;
    tfloat3* defforce; tfloat3* force;
    // engine: cos and sin of th0 per edge (NULL: computed from th0)
    const float* cth0 = nullptr; const float* sth0 = nullptr;
    // engine only (NULL in the entry point): Groot-Warren corrector after
    // the membrane force, vout = (velhalf + dt/2 F/m, rho of velrho)
    const tfloat3* velhalf = nullptr; const float4* velrho = nullptr; float4* vout = nullptr;
    // engine only: unwrapped vertex positions in id order, x[nm], y[nm],
    // z[nm] (FP64); replaces the coor/nthro/curpos gathers of the vertices
    const double* mpos = nullptr; int nm = 0;
    // batched runs: per-vertex replica (physics of the vertex's suspension)
    const GcRep* reps = nullptr; const int* cell_rep = nullptr; const int* celltab = nullptr;
    // engine: fault word (GC_FAULT_* bits OR-ed in on a non-finite vertex
    // force or an edge at or past l_max); NULL: not watched
    int* fault = nullptr;
};

// reps/cell_rep: batched runs (NULL for one suspension) -- the cell's
// replica supplies the parameter block and the box.
__global__ void k_geo_cell(GcParams P, long stepcount, const GcRep* __restrict__ reps,
                           const int* __restrict__ cell_rep, const float* __restrict__ bbox,
                           const int* __restrict__ curpos, const tint3* __restrict__ tri,
                           const tint2* __restrict__ betri, const tint3* __restrict__ nthro,
                           const tfloat3* __restrict__ coor, tfloat3* __restrict__ trinor,
                           float* __restrict__ triarea, float* __restrict__ reftriarea,
                           tfloat3* __restrict__ tricen, float* __restrict__ refarea_para,
                           float* __restrict__ refvol_para, float* __restrict__ area_para,
                           float* __restrict__ vol_para, float* __restrict__ area,
                           float* __restrict__ vol);
__global__ void k_par_area(int nm, const int* __restrict__ ts, const int* __restrict__ te,
                           const float* __restrict__ triarea, float* __restrict__ pararea);
__global__ void k_mech_para(GcParams P, int nedge, const int6* __restrict__ edge,
                            const int* __restrict__ curpos, const tfloat3* __restrict__ coor,
                            float* __restrict__ l0o, float* __restrict__ lmaxo,
                            const float* __restrict__ kc, const float* __restrict__ mu0,
                            float* __restrict__ kbo, float* __restrict__ mo, float* __restrict__ po,
                            float* __restrict__ kpo, float* __restrict__ th0o);
__global__ void k_def_force(GcParams P, int nm, const float* __restrict__ bbox, DefArgs a);
// Fused per-cell membrane step (one 1024-thread workgroup per cell): the
// vertex positions, then k_geo_cell's geometry and cell area/volume, then
// the edge and triangle forces held in LDS and summed per vertex in the
// incidence order, as the three kernels below do through global memory.
// A cell fits when its vertices, edges and triangles are contiguous, its
// records reference only its own vertices, and it has at most GC_CELL_NV
// vertices, GC_CELL_NE edges and GC_CELL_NT triangles.
constexpr int GC_CELL_BS = 1024;
constexpr int GC_CELL_NV = 768;
constexpr int GC_CELL_NE = 2048;
constexpr int GC_CELL_NT = (GC_CELL_NE * 12) / 9;
constexpr int GC_GEO_BS = GC_CELL_BS;   // threads of k_geo_cell (sets its FP64 reduction order)
constexpr int GC_CELL_INC = 12;   // edge incidences per vertex read together (more are read one by one)
constexpr int GC_CELL_TINC = 8;   // triangle incidences per vertex read together
struct CellArgs {
    const int4* range;    // (first vertex, vertices, first edge, edges) per cell, 0-based
    const tint2* betri;   // triangle range per cell (inclusive)
    long step;            // -1: also the reference geometry (k_geo_cell's stepcount)
    int round_geo;        // geometry from FP32-rounded positions (membrane substeps)
    tfloat3* trinor; float* triarea; float* reftriarea; tfloat3* tricen;
    float* refarea_para; float* refvol_para; float* area_para; float* vol_para; float* area; float* vol;
    const int* gate = nullptr;   // non-null: run only when *gate != 0 (device-decided fallback)
};
__global__ void k_cell_membrane(GcParams P, const float* __restrict__ bbox, CellArgs ca, DefArgs a);
// c[i], s[i] = cos, sin of th[i] (per-edge reference dihedral angle)
__global__ void k_sincos(int n, const float* __restrict__ th, float* __restrict__ c, float* __restrict__ s);

// Membrane substepping (velocity Verlet of the membrane under a fixed fluid
// force): the kick and the drift, shared by the per-substep kernels and the
// fused per-cell kernel so both give the same bits.
__device__ __forceinline__ void ms_kick(float3& u, const float3& fe, const float3& fm, float dts, float inv_m)
{
    const float h = 0.5f * dts * inv_m;
    u.x += h * (fe.x + fm.x); u.y += h * (fe.y + fm.y); u.z += h * (fe.z + fm.z);
}
__device__ __forceinline__ double ms_drift(double x, float dts, float u) { return x + (double)dts * u; }
// State of the membrane system in id order (nm = DefArgs::nm vertices):
// positions x[nm], y[nm], z[nm] (FP64), velocity, fluid force (fixed over
// the step), membrane force, FP32 positions; mx6/mx7 (or NULL) receive the
// positions after substeps k-2 and k-1.
struct SubArgs {
    int k; float dts, inv_m;
    const double* mx_in; const tfloat3* mv_in; const tfloat3* fm_in;   // start state (may be mx, mv, fm)
    double* mx; tfloat3* mv; const tfloat3* fe; tfloat3* fm; tfloat3* mcoor;   // end state
    // non-null: positions after substeps k-2 and k-1 into mx6 / mx7 and
    // d2x = (x_k - x_{k-1}) - (x_{k-1} - x_{k-2}) (FP64, then FP32)
    double* mx6; double* mx7; tfloat3* d2x;
};
// all k substeps of each cell in one workgroup (cells as k_cell_membrane;
// ca.step must not be -1)
__global__ void k_cell_substeps(GcParams P, CellArgs ca, DefArgs a, SubArgs sa);
// host check of the conditions above; fills range (C entries)
bool cell_ranges(int nm, int nedge, int ntri, int ncell, const int* celltab, const tint2* betri,
                 const int6* edge, const tint3* tri, int4* range);

// the same force in three passes: per edge (ef[4 e + role]), per triangle
// (tf[3 t + corner]), then the per-vertex sum in incidence order
__global__ void k_edge_force(GcParams P, int nedge, const float* __restrict__ bbox, DefArgs a,
                             tfloat3* __restrict__ ef);
__global__ void k_tri_force(GcParams P, int ntri, const float* __restrict__ bbox, DefArgs a,
                            tfloat3* __restrict__ tf);
__global__ void k_vert_force(GcParams P, int nm, DefArgs a, const tfloat3* __restrict__ ef,
                             const tfloat3* __restrict__ tf);
__global__ void k_agg_force(int n, GcParams P, const tfloat3* __restrict__ coor,
                            tfloat3* __restrict__ aggforce, const tint3* __restrict__ cc,
                            const int* __restrict__ ncube, const int* __restrict__ sortpart,
                            const int* __restrict__ celltab, const float* __restrict__ outer,
                            const tint2* __restrict__ be, tfloat3* __restrict__ force);
int build_incidence(int nm, int nedge, const int6* h_edge, int ntri, const tint3* h_tri,
                    std::vector<int>& es, std::vector<int>& ee, std::vector<int>& ts,
                    std::vector<int>& te);

// gc_abi.hip
__global__ void k_fill_int(int* __restrict__ a, int v, int n);

}  // namespace gc
