// gc_engine.hip -- the engine: library-owned state and the step of
// Particle::UpdatePar (src/Particle.cu:1310-1329) as a fixed sequence of
// fused kernels on one stream, captured into a HIP graph for replay.  One
// engine holds one suspension (gc_EngineCreate) or a batch of independent
// suspensions of any sizes stepped together by the same kernels
// (gc_EnsembleCreate).
//
// Per step (reference launch numbers of UpdatePar in brackets):
//   k2_integrate   [1,2,3]  r + dt v + dt^2/2 F/m in FP64 and periodic wrap
//                           (not stored), cube key, histogram rank (atomic)
//   k2_tile_sums,  [4]      cube offsets: tile totals, then each tile's scan
//   k2_cube_offsets         with the preceding totals; noise counter advance
//   k2_scatter     [4]      (cube, id, old slot) into cube segments
//   k2_sortgather  [1,2,4,5,6,8] rank within the cube by id -> final slot;
//                           gather of the state advanced as in the key pass;
//                           Groot-Warren predictor
//   k2_density     [7,9]    number density (moving) and no-slip wall
//                           velocity (Bound2) in one neighbour pass; writes
//                           each moving particle's neighbour list
//   k2_force       [10-13,17] EOS on the fly, SDPD pair force, Bound1
//                           repulsion, body force over that list; corrector
//                           for fluid
//   k2_agg         [16]     Morse between cells (when De != 0), one thread
//                           per membrane vertex
//   k_cell_membrane [14,15,17] per cell: geometry, FP64 cell area/volume,
//                           membrane forces, corrector for membrane (the
//                           three-pass form k_geo_cell + k2_mgather +
//                           k_edge/tri/vert_force for cells that do not fit)
//   k2_wallvel     [18]     wall velocity from the corrected flow (on export)
//
// Layouts (slot order): posid float4 (x, y, z, id bits), velrho float4
// (predicted v, rho) read by neighbour loops, vout float4 (v(t+dt), rho) =
// the reference's Velrhop after the step, pos64 = x[N], y[N], z[N] doubles
// (accumulator), coor/force/velhalf in the TypesDef float3 layout.
// BeginEndCube is the scan's offsets array: cube k holds [offs[k], offs[k+1]).
//
// Batches: replicas are concatenated with the id layout [Bound1 of all |
// Bound2 of all | fluid of all | membrane of all] and disjoint ranges of cube
// keys; a kernel finds a particle's replica from its id (rep_of) and reads
// that replica's GcRep (parameters, grid, box).  Slots of one replica are
// contiguous and ordered by (cube, id) as when it runs alone, and the pair
// noise uses the replica's own ids and seed, so every replica evolves bit
// for bit as its solo run.
#include "gc_engine_state.h"

#include <algorithm>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <type_traits>
#include <vector>

using namespace gc;

namespace {

constexpr int BS = GC_BS;
// Neighbour list written by the density pass for every moving particle:
// entry k of slot p at nl[k * N + p] (column-major: a wave's lanes write and
// read consecutive words), value (slot << 6) | image code.  A particle with
// more than NL_CAP neighbours gets count -1 and the force pass scans its
// cubes instead.
constexpr int NL_CAP = 64;
// Skin mode: Verlet lists of up to VL_CAP entries, and after them one row of
// neighbour bits per 32 entries (density pass -> force pass): VL_ROWS rows.
constexpr int VL_CAP = 128, VL_ROWS = VL_CAP + VL_CAP / 32;
// ints per wave block of the neighbour lists
inline int nl_stride(const gc_engine* e) { return (e->skin_mode() ? VL_ROWS : NL_CAP) * 64; }
// Skin mode: grid of the gated rebuild kernels (they stride over their
// logical groups; an empty launch then dispatches at most this many)
constexpr int GATED_MAX_GROUPS = 2048;
inline int gated_grid(int nb) { return std::min(xcd_grid(nb), GATED_MAX_GROUPS); }
// A batch rebuilds a few replicas per step: half the grid (64 sheared
// replicas: rebuild 252 -> 239 us per step; a full 10M rebuild is faster
// with the whole grid, 126 against 130 us per step amortised).
inline int gated_grid(int nb, bool batch) { return std::min(xcd_grid(nb), batch ? GATED_MAX_GROUPS / 2 : GATED_MAX_GROUPS); }
// a batch's rebuild: activity bytes per 64-wide chunk of slots (then of
// cubes), each table padded to whole 32-bit words
inline int chunks64(int n) { return (n + 63) / 64; }
inline size_t act_bytes(int n) { return ((size_t)chunks64(n) + 3) / 4 * 4; }
// Skin mode: workgroup displacement maxima, the advance's groups (at most
// one per 64 particles) first, the membrane substeps' groups after them.
inline int skin_ms_off(const gc_engine* e) { return (e->N + 63) / 64; }
inline size_t skin_bmax_size(const gc_engine* e)
{
    return DSTAT_N * ((size_t)skin_ms_off(e) + (e->nm + GC_BS - 1) / GC_BS + 1);
}
constexpr int CUBE_SCAN_BS = 256, CUBE_IPT = 4, CUBE_TILE = CUBE_SCAN_BS * CUBE_IPT;

// Image code of a range: 2 bits per axis, 1 = shift -L, 2 = shift +L.
__device__ __forceinline__ int shift_code(float sx, float sy, float sz)
{
    return (sx < 0.f ? 1 : (sx > 0.f ? 2 : 0)) | ((sy < 0.f ? 1 : (sy > 0.f ? 2 : 0)) << 2) |
           ((sz < 0.f ? 1 : (sz > 0.f ? 2 : 0)) << 4);
}

__device__ __forceinline__ float code_shift(int c, float L)
{
    return (c & 3) == 1 ? -L : ((c & 3) == 2 ? L : 0.0f);
}

// posid.w = id | (wrap code << ID_BITS).  In skin mode the wrap code holds a
// particle's periodic wraps since the last list rebuild (an image code: per
// axis 1 = -1, 2 = +1); it is 0 in per-step mode and right after a rebuild.
// Ids are below 2^25.
constexpr int ID_BITS = 25;
constexpr int ID_MASK = (1 << ID_BITS) - 1;
__device__ __forceinline__ int pid_of(float w) { return __float_as_int(w) & ID_MASK; }
__device__ __forceinline__ int woff_of(float w) { return (int)((unsigned)__float_as_int(w) >> ID_BITS); }
__device__ __forceinline__ float id_word(int id, int woff) { return __int_as_float(id | (woff << ID_BITS)); }

// L coefficient (-1, 0, +1) of axis a in an image code
__device__ __forceinline__ int code_coef(int c, int a)
{
    const int b = (c >> (2 * a)) & 3;
    return (b & 2) ? 1 : -(b & 1);
}
__device__ __forceinline__ int coef_code(int k0, int k1, int k2)
{
    auto e = [](int k) { return k < 0 ? 1 : (k > 0 ? 2 : 0); };
    return e(k0) | (e(k1) << 2) | (e(k2) << 4);
}

// Skin mode: the image shift of a list entry along axis a.  The entry's code
// was fixed at the rebuild; positions are wrapped every step, so a particle
// that wrapped since (wrap code o) sits o L away from where the code
// assumed: S + (o_j - o_i) L, an exact multiple of L (the physical image of
// the pair, i.e. the shift the per-step cube scan would use).
__device__ __forceinline__ float eff_shift(int code, int oi, int oj, int a, float L)
{
    const int k = code_coef(code, a) + code_coef(oj, a) - code_coef(oi, a);
    return (float)k * L;
}

// Slots [jb, je) in order, four loads issued before their use.
template <class Fn>
__device__ __forceinline__ void scan_range(const float4* __restrict__ posid, int jb, int je, Fn&& fn)
{
    int j = jb;
    for (; j + 4 <= je; j += 4) {
        float4 a = posid[j], b = posid[j + 1], c = posid[j + 2], d = posid[j + 3];
        fn(j, a); fn(j + 1, b); fn(j + 2, c); fn(j + 3, d);
    }
    for (; j < je; ++j) fn(j, posid[j]);
}

// Logical groups [0, nb) of TBS lanes over a grid of gridDim.x groups (a
// multiple of 8, renumbered so each XCD takes contiguous runs): body(lb) for
// lb = xcd_block(gridDim.x) + k gridDim.x.  With gridDim.x = xcd_grid(nb)
// each group runs one logical group, numbered as xcd_block(nb) does; the
// skin mode's gated rebuild kernels run on a small fixed grid, so that an
// empty launch (no rebuild this step) dispatches few groups.
template <class Body>
__device__ __forceinline__ void for_groups(int nb, Body&& body)
{
    const int G = gridDim.x;   // the renumbering is a permutation for multiples of 8 only
    for (int lb = G % GC_NXCD == 0 ? xcd_block(G) : (int)blockIdx.x; lb < nb; lb += G) body(lb);
}

// Skin mode, a batch: replica r holds slots [sb[r], sb[r + 1]) (and, with
// the cube bases, cubes [cb[r], cb[r + 1])).  The replica of slot (or cube)
// x, by bisection over the nrep + 1 bases.
__device__ __forceinline__ int base_rep(const int* __restrict__ sb, int nrep, int x)
{
    int lo = 0, hi = nrep;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (sb[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Activity of a group of TBS slots (or cubes) from the bytes of its 64-wide
// chunks (1: a rebuilding replica owns part of the chunk), one load.
template <int TBS>
__device__ __forceinline__ bool group_active(const unsigned char* __restrict__ act, int lb)
{
    static_assert(TBS == 64 || TBS == 128 || TBS == 256, "group of 1, 2 or 4 chunks");
    if constexpr (TBS == 256) return reinterpret_cast<const unsigned*>(act)[lb] != 0u;
    else if constexpr (TBS == 128) return reinterpret_cast<const unsigned short*>(act)[lb] != 0;
    else return act[lb] != 0;
}

// for_groups over the active groups only (a batch's rebuild: the groups of
// the rebuilding replicas).  The activity of all of a workgroup's logical
// groups is loaded at once, one per lane, so an idle group costs no load of
// its own; act == nullptr: every group.
template <int TBS, class Body>
__device__ __forceinline__ void for_groups_of(int nb, const unsigned char* __restrict__ act, Body&& body)
{
    if (!act) { for_groups(nb, body); return; }
    const int G = gridDim.x;
    const int first = G % GC_NXCD == 0 ? xcd_block(G) : (int)blockIdx.x;
    const int lane = threadIdx.x & 63;
    for (int base = first; base < nb; base += 64 * G) {
        const int lb = base + lane * G;
        unsigned long long m = __ballot(lb < nb && group_active<TBS>(act, lb));
        while (m) {
            const int k = (int)__ffsll((long long)m) - 1;
            m &= m - 1;
            body(base + k * G);
        }
    }
}

__device__ __forceinline__ void key_to_cc(int k, const GcRep& R, int& cx, int& cy, int& cz)
{
    cube_coords(R, k, cx, cy, cz);
}

// Runs body(R) with the lane's replica record.  A batch wave whose lanes
// all belong to one replica (all but the few waves at replica boundaries)
// passes that replica's record at a wave-uniform address, so its fields
// are read with scalar loads once per wave instead of vector loads per lane;
// body is inlined for both cases.  live: lanes with a valid id.
template <bool ENS, class Body>
__device__ __forceinline__ void with_rep(const GcRep& R0, const GcRep* __restrict__ reps,
                                         const int* __restrict__ rep_of, bool live, int id, Body&& body)
{
    if constexpr (ENS) {
        const int r = live ? rep_of[id] : -1;
        const int r0 = __builtin_amdgcn_readfirstlane(r);
        if (__ballot(live && r != r0) == 0) body(reps[r0]);
        else body(reps[r < 0 ? 0 : r]);
    } else {
        body(R0);
    }
}

// The replica of global id `id`: the kernel argument for one suspension,
// the table entry for a batch (a reference either way: a by-value copy of
// the whole record costs registers or scratch).
template <bool ENS>
__device__ __forceinline__ const GcRep& rep_of_id(const GcRep& R0, const GcRep* __restrict__ reps,
                                                  const int* __restrict__ rep_of, int id)
{
    if constexpr (ENS) return reps[rep_of[id]];
    else return R0;
}

// Contiguous slot ranges of the 27 neighbour cubes (x-fastest keys: one
// range per non-wrapping (y, z) row, plus the periodic image cube on a
// wrapping x side), with the image shift of each range.  Order per row:
// image at -x, image at +x, main range.  fn has a single call site, so a
// heavy body is inlined once.
template <class Fn>
__device__ __forceinline__ void visit2(const GcRep& R, const int* __restrict__ offs_all, int cx,
                                       int cy, int cz, Fn&& fn)
{
    const int* offs = offs_all + R.cube_off;
    const int* per = R.P.per;
    const int nx = R.nc[0];
    const bool wlo = per[0] && cx == 0, whi = per[0] && cx == nx - 1;
    const int xlo = cx > 0 ? cx - 1 : 0, xhi = cx < nx - 1 ? cx + 1 : nx - 1;
    const int np = 1 + (int)wlo + (int)whi;
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz) {
        int z = cz + dz;
        float sz = 0.0f;
        if (z < 0) { if (!per[2]) continue; z += R.nc[2]; sz = -R.L[2]; }
        else if (z >= R.nc[2]) { if (!per[2]) continue; z -= R.nc[2]; sz = R.L[2]; }
#pragma unroll 1
        for (int dy = -1; dy <= 1; ++dy) {
            int y = cy + dy;
            float sy = 0.0f;
            if (y < 0) { if (!per[1]) continue; y += R.nc[1]; sy = -R.L[1]; }
            else if (y >= R.nc[1]) { if (!per[1]) continue; y -= R.nc[1]; sy = R.L[1]; }
            const int row = cube_row(R, y, z) * nx;
#pragma unroll 1
            for (int k = 0; k < np; ++k) {
                // piece k: the -x image (if wlo), the +x image (if whi), the main range
                bool img_lo = wlo && k == 0;
                bool img_hi = whi && k == (int)wlo;
                int a = img_lo ? nx - 1 : (img_hi ? 0 : xlo);
                int b = img_lo ? nx : (img_hi ? 1 : xhi + 1);
                float sx = img_lo ? -R.L[0] : (img_hi ? R.L[0] : 0.0f);
                fn(offs[row + a], offs[row + b], sx, sy, sz);
            }
        }
    }
}

__device__ __forceinline__ float ext_force_x(const GcParams& P, float y)
{
    float f = P.fext;
    return (P.force_mode == GC_FORCE_REVERSE_POISEUILLE && y >= P.force_split) ? -f : f;
}

// [1,2] r += dt v + dt^2/(2m) F in FP64 and the periodic wrap (image
// counters in img).  The key pass and the gather both run it on the same
// inputs, so the gather reproduces the key pass's positions bit for bit
// and the advanced positions are not stored in between.
// One periodic wrap of u (a position at most one period outside the box).
__device__ __forceinline__ void wrap_once(const GcRep& R, const GcParams& P, double u[3], int img[3])
{
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!P.per[d]) continue;
        double lo = R.bbox[2 * d], hi = R.bbox[2 * d + 1];
        double L = (double)(R.bbox[2 * d + 1] - R.bbox[2 * d]);
        if (u[d] < lo) { u[d] += L; img[d] -= 1; }
        else if (u[d] >= hi) { u[d] -= L; img[d] += 1; }
    }
}

// r + dt v + dt^2/(2m) F in FP64, not wrapped
__device__ __forceinline__ void advance_step(const GcParams& P, const float4& v, const tfloat3& f, double u[3])
{
    double dt = P.dt, a = 0.5 * (double)P.dt * (double)P.dt * (double)P.inv_m;
    u[0] = u[0] + (dt * v.x + a * f.x);
    u[1] = u[1] + (dt * v.y + a * f.y);
    u[2] = u[2] + (dt * v.z + a * f.z);
}

__device__ __forceinline__ void advance(const GcRep& R, const GcParams& P, const float4& v, const tfloat3& f,
                                        double& x, double& y, double& z, int img[3])
{
    double u[3] = {x, y, z};
    advance_step(P, v, f, u);
    wrap_once(R, P, u, img);
    x = u[0]; y = u[1]; z = u[2];
}

// A wall's entry of velrho before the density pass: density rho0 (walls
// keep it) and no velocity (Bound2 walls get theirs in the density pass).
// The walls' vout entries are not carried through the re-sort (the force
// pass writes vout only for moving particles, at their new slots), so after
// the slots shift a wall's old slot holds another particle's vout: it is
// not read here.
template <bool ENS>
__device__ __forceinline__ float4 wall_velrho(const GcRep& R0, const GcRep* __restrict__ reps,
                                              const int* __restrict__ rep_of, int id)
{
    const float rho0 = ENS ? reps[rep_of[id]].P.rho0 : R0.P.rho0;
    return make_float4(0.f, 0.f, 0.f, rho0);
}

__device__ __forceinline__ bool moves(const GcParams& P, int id, int mmove)
{
    return id >= P.nf && (mmove || id < P.nc);
}

// Moving Bound2 walls (plane Couette shear) translate along x at their
// wall velocity; with wmove = 0 (no wall of any suspension moves) they stay.
__device__ __forceinline__ bool wall_moves(const GcParams& P, int id, int wmove)
{
    return wmove && id >= P.nb2 && id < P.nf;
}
__device__ __forceinline__ float4 wall_v4(const GcParams& P, double y)
{
    return make_float4((float)y >= P.wall_split ? P.wall_u_hi : P.wall_u_lo, 0.f, 0.f, 0.f);
}

// Histogram rank of key k for the wave's active lanes (a prefix of the
// wave): each run of equal keys takes its ranks with one atomic.
__device__ __forceinline__ void hist_rank(bool act, int k, int p, int* __restrict__ rank, int* __restrict__ cnt)
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(k, 1);
    // runs of equal keys; an inactive lane ends a run (the active lanes
    // need not be a prefix of the wave)
    const bool brk = !act || lane == 0 || prev != k;
    const unsigned long long heads = __ballot(brk);
    const unsigned long long above = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
    const int next = above ? (int)__ffsll((long long)above) - 1 : 64;
    int base = 0;
    if (act && brk) base = atomicAdd(&cnt[k], next - lane);
    // the nearest break at or below an active lane is its run's head
    const unsigned long long upto = lane == 63 ? heads : heads & ((2ull << lane) - 1ull);
    const int myhead = 63 - __clzll((long long)upto);   // lane 0 always breaks
    base = __shfl(base, myhead);
    if (act) rank[p] = base + (lane - myhead);
}

// [1,2,3] key of the advanced position, histogram rank.  Slots are in the
// previous step's (cube, id) order and few particles change cube per step,
// so consecutive lanes mostly share a key: each run of equal keys in a wave
// takes its ranks with one atomic (base + offset in the run), about a third
// of the device-scope atomics of one per particle.
template <bool ENS, int TBS>
__global__ __launch_bounds__(TBS) void k2_integrate(int n, GcRep R0, const GcRep* __restrict__ reps,
    const int* __restrict__ rep_of, int move, int mmove, int wmove, const double* __restrict__ pos64,
    const float4* __restrict__ vout, const tfloat3* __restrict__ force, const int* __restrict__ sp,
    int* __restrict__ key, int* __restrict__ rank, int* __restrict__ cnt)
{
    const int p = xcd_block((n + TBS - 1) / TBS) * TBS + threadIdx.x;
    const bool act = p < n;
    if (__ballot(act) == 0) return;   // wave-uniform
    int k = -1;
    const int id = act ? sp[p] : 0;
    with_rep<ENS>(R0, reps, rep_of, act, id, [&](const GcRep& R) __attribute__((always_inline)) {
    if (act) {
        const GcParams& P = R.P;
        double x = pos64[p], y = pos64[n + p], z = pos64[2 * n + p];
        if (move && moves(P, id, mmove)) {
            int img[3] = {0, 0, 0};
            advance(R, P, vout[p], force[p], x, y, z, img);
        } else if (move && wall_moves(P, id, wmove)) {
            int img[3] = {0, 0, 0};
            advance(R, P, wall_v4(P, y), tfloat3{0.f, 0.f, 0.f}, x, y, z, img);
        }
        int cx = cube_coord((float)x, R.bbox[0], P.off[0], R.invw[0], R.nc[0]);
        int cy = cube_coord((float)y, R.bbox[2], P.off[1], R.invw[1], R.nc[1]);
        int cz = cube_coord((float)z, R.bbox[4], P.off[2], R.invw[2], R.nc[2]);
        k = R.cube_off + cx + R.nc[0] * cube_row(R, cy, cz);
        key[p] = k;
    }
    });
    hist_rank(act, k, p, rank, cnt);
}

// [4] cube counts per tile of CUBE_TILE cubes
__global__ __launch_bounds__(CUBE_SCAN_BS) void k2_tile_sums(int K, const int* __restrict__ cnt,
                                                             int* __restrict__ tsum, const int* __restrict__ gate)
{
    if (gate && *gate == 0) return;   // skin mode: no rebuild this step
    __shared__ int red[CUBE_SCAN_BS];
    const int tid = threadIdx.x, base = blockIdx.x * CUBE_TILE + tid * CUBE_IPT;
    int t = 0;
#pragma unroll
    for (int k = 0; k < CUBE_IPT; ++k) t += base + k < K ? cnt[base + k] : 0;
    red[tid] = t;
    __syncthreads();
    for (int o = CUBE_SCAN_BS / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) tsum[blockIdx.x] = red[0];
}

// [4] BeginEndCube (exclusive offsets of the cube counts, total at K): block
// b sums the totals of the tiles before it and scans its own tile.  In a
// step it also advances the noise counter (the force pass of this step
// reads the value before the advance).
__global__ __launch_bounds__(CUBE_SCAN_BS) void k2_cube_offsets(int K, const int* __restrict__ cnt,
    const int* __restrict__ tsum, int* __restrict__ offs, unsigned long long* __restrict__ ctr, int bump,
    int* __restrict__ nlover, const int* __restrict__ gate)
{
    if (gate && *gate == 0) return;
    __shared__ int red[CUBE_SCAN_BS];
    const int b = blockIdx.x, tid = threadIdx.x;
    int pre = 0;
    for (int i = tid; i < b; i += CUBE_SCAN_BS) pre += tsum[i];
    red[tid] = pre;
    __syncthreads();
    for (int o = CUBE_SCAN_BS / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    const int before = red[0];
    __syncthreads();
    const int base = b * CUBE_TILE + tid * CUBE_IPT;
    int v[CUBE_IPT];
    int t = 0;
#pragma unroll
    for (int k = 0; k < CUBE_IPT; ++k) {
        v[k] = base + k < K ? cnt[base + k] : 0;
        t += v[k];
    }
    red[tid] = t;
    __syncthreads();
    for (int o = 1; o < CUBE_SCAN_BS; o <<= 1) {
        int a = tid >= o ? red[tid - o] : 0;
        __syncthreads();
        red[tid] += a;
        __syncthreads();
    }
    int run = before + red[tid] - t;
#pragma unroll
    for (int k = 0; k < CUBE_IPT; ++k) {
        if (base + k < K) offs[base + k] = run;
        run += v[k];
    }
    if (b == (int)gridDim.x - 1 && tid == CUBE_SCAN_BS - 1) offs[K] = before + red[CUBE_SCAN_BS - 1];
    if (b == 0 && tid == 0) {
        if (bump) *ctr += 1ull;
        *nlover = 0;   // list overflows of this step's density pass
    }
}

// [4] (cube, id) into cube segments; clears the histogram for the next step
// (every non-empty cube has exactly one rank-0 member).
// Skin mode, a batch (rbflag set): the particles of replicas that do not
// rebuild keep their slots and are not scattered, and the counts persist
// (cleared per replica before its key pass).
__global__ __launch_bounds__(BS) void k2_scatter(int n, const int* __restrict__ key,
    const int* __restrict__ rank, const int* __restrict__ sp, const int* __restrict__ offs,
    int* __restrict__ tk, int* __restrict__ tv, int* __restrict__ to, int* __restrict__ cnt,
    const int* __restrict__ gate, const int* __restrict__ rep_of, const int* __restrict__ rbflag,
    const unsigned char* __restrict__ gact)
{
    if (gate && *gate == 0) return;
    for_groups_of<BS>((n + BS - 1) / BS, gact, [&](int lb) {
    int p = lb * BS + threadIdx.x;
    if (p >= n) return;
    const int id = sp[p];
    if (rbflag && !rbflag[rep_of[id]]) return;
    int k = key[p], r = rank[p];
    if (!rbflag && r == 0) cnt[k] = 0;
    int q = offs[k] + r;
    if (q < 0 || q >= n) return;   // only if the counts were inconsistent
    tk[q] = k;
    tv[q] = id;
    to[q] = p;
    });
}

// [1,2,4,5,6,8] final slot = cube start + number of smaller ids in the
// cube; gather of the particle from its old slot (carried by the scatter),
// advanced as in the key pass (move); predictor.  CurPos is kept for the
// membrane vertices only (the membrane kernels read it; the whole map is
// the inverse of SortPart).
// SK (skin mode, a list rebuild): the positions were advanced and wrapped
// in place by the advance pass; the wrap codes of a rebuilding replica are
// cleared (its lists start from these positions), the slots of the other
// replicas of a batch are not touched; the predictor is recomputed from the
// same inputs as in the advance pass (the same bits).
template <bool ENS, bool SK>
__global__ __launch_bounds__(BS) void k2_sortgather(int n, GcRep R0, const GcRep* __restrict__ reps,
    const int* __restrict__ rep_of, int move, int mmove, int wmove, const int* __restrict__ tk,
    const int* __restrict__ tv, const int* __restrict__ to, const int* __restrict__ offs,
    int* __restrict__ curpos,
    const double* __restrict__ pos_old, double* __restrict__ pos_new,
    const float4* __restrict__ vout, const tfloat3* __restrict__ force,
    const tint3* __restrict__ nt_old, tint3* __restrict__ nt_new, int* __restrict__ sp_new,
    int* __restrict__ key_new, float4* __restrict__ posid, tfloat3* __restrict__ coor,
    float4* __restrict__ velrho, tfloat3* __restrict__ velhalf, const int* __restrict__ wincl,
    int* __restrict__ act, int* __restrict__ fault, const int* __restrict__ gate,
    const int* __restrict__ rbflag, unsigned char* __restrict__ woff, const int* __restrict__ sp_old,
    const unsigned char* __restrict__ gact)
{
    if (SK && *gate == 0) return;   // skin mode: no rebuild this step
    for_groups_of<BS>((n + BS - 1) / BS, SK && ENS ? gact : nullptr, [&](int lb) {
    int q = lb * BS + threadIdx.x;
    if (q >= n) return;
    // a batch: the slots of a replica that does not rebuild are left as the
    // advance pass wrote them (a replica's slot range is fixed, as its
    // particle count is)
    if (SK && ENS && !rbflag[rep_of[sp_old[q]]]) return;
    int k = tk[q], id = tv[q];
    int b = max(offs[k], 0), e = min(offs[k + 1], n);
    int r = 0;
    for (int t = b; t < e; ++t) r += tv[t] < id;
    int s = b + r;
    int o = to[q];
    double x = pos_old[o], y = pos_old[n + o], z = pos_old[2 * n + o];
    tint3 nt = nt_old[o];
    float4 v = vout[o];
    const bool mv = move && moves(R0.P, id, mmove);
    const bool wmv = move && wall_moves(R0.P, id, wmove);
    with_rep<ENS>(R0, reps, rep_of, SK || mv || wmv, id, [&](const GcRep& R) __attribute__((always_inline)) {
    if constexpr (SK) {
        const GcParams& P = R.P;
        if (moves(P, id, mmove)) {
            const tfloat3 f = force[o];
            float h = 0.5f * P.dt * P.inv_m, l = P.lam * P.dt * P.inv_m;
            velhalf[s] = tfloat3{v.x + h * f.x, v.y + h * f.y, v.z + h * f.z};
            v.x += l * f.x; v.y += l * f.y; v.z += l * f.z;
        } else {
            velhalf[s] = id < P.nf ? tfloat3{0.f, 0.f, 0.f} : tfloat3{v.x, v.y, v.z};
        }
    } else {
    if (mv) {
        const GcParams& P = R.P;
        const tfloat3 f = force[o];
        int img[3] = {0, 0, 0};
        advance(R, P, v, f, x, y, z, img);
        nt.x += img[0]; nt.y += img[1]; nt.z += img[2];
        float h = 0.5f * P.dt * P.inv_m, l = P.lam * P.dt * P.inv_m;
        velhalf[s] = tfloat3{v.x + h * f.x, v.y + h * f.y, v.z + h * f.z};
        v.x += l * f.x; v.y += l * f.y; v.z += l * f.z;
    } else {
        // walls: half-step velocity zero (independent of the exported wall
        // velocity, so reading the state changes nothing)
        velhalf[s] = id < R0.P.nf ? tfloat3{0.f, 0.f, 0.f} : tfloat3{v.x, v.y, v.z};
        if (wmv) {   // a moving wall: the key pass's advance, the same bits
            int img[3] = {0, 0, 0};
            advance(R, R.P, wall_v4(R.P, y), tfloat3{0.f, 0.f, 0.f}, x, y, z, img);
            nt.x += img[0]; nt.y += img[1]; nt.z += img[2];
        }
    }
    }
    });
    pos_new[s] = x; pos_new[n + s] = y; pos_new[2 * n + s] = z;
    // a non-finite position has no neighbours (every distance test fails)
    // and so a finite force: it is caught here
    if (!isfinite(x + y + z)) atomicOr(fault, GC_FAULT_NONFINITE);
    float fx = (float)x, fy = (float)y, fz = (float)z;
    if (SK) woff[s] = 0;   // the new lists start from these positions
    posid[s] = make_float4(fx, fy, fz, __int_as_float(id));
    // skin mode: coor is read for membrane vertices only (posid holds the rest)
    if (!SK || id >= R0.P.nc) coor[s] = tfloat3{fx, fy, fz};
    nt_new[s] = nt;
    sp_new[s] = id;
    key_new[s] = k;
    velrho[s] = id < R0.P.nf ? wall_velrho<ENS>(R0, reps, rep_of, id) : v;
    if (id >= R0.P.nc) curpos[id] = s;
    if (!wincl) {
        act[s] = s;   // moving walls: the force pass runs over every slot
    } else if (id >= R0.P.nf) {
        // a cube holds its walls (the smallest ids) first, so the moving
        // particles before this one are s minus the walls of cubes <= k
        const int ai = s - wincl[k];
        if ((unsigned)ai < (unsigned)n) act[ai] = s;
    }
    });
}

// [7,9] density of moving particles; velocity of Bound2 walls from the
// predicted velocities of their moving neighbours.
// One (y, z) row of a lane's 27-cube neighbourhood: the main x range
// [ma, mb) (no x shift) and, for a cube on a periodic x face, the image
// cube [wa, wb) on the far face with x shift wsx; (sy, sz) the row's image
// shift.  ok = false for a row beyond a non-periodic face.
struct RowRanges {
    int ma, mb, wa, wb;
    float wsx, sy, sz;
    bool ok;
};

__device__ __forceinline__ RowRanges row_ranges(const GcRep& R, const int* __restrict__ offs_all, int cx,
                                                int cy, int cz, int q)
{
    RowRanges r;
    r.ok = false; r.ma = r.mb = r.wa = r.wb = 0; r.wsx = r.sy = r.sz = 0.0f;
    const int* offs = offs_all + R.cube_off;
    const int* per = R.P.per;
    int z = cz + q / 3 - 1, y = cy + q % 3 - 1;
    if (z < 0) { if (!per[2]) return r; z += R.nc[2]; r.sz = -R.L[2]; }
    else if (z >= R.nc[2]) { if (!per[2]) return r; z -= R.nc[2]; r.sz = R.L[2]; }
    if (y < 0) { if (!per[1]) return r; y += R.nc[1]; r.sy = -R.L[1]; }
    else if (y >= R.nc[1]) { if (!per[1]) return r; y -= R.nc[1]; r.sy = R.L[1]; }
    const int nx = R.nc[0], row = cube_row(R, y, z) * nx;
    const int xlo = cx > 0 ? cx - 1 : 0, xhi = cx < nx - 1 ? cx + 1 : nx - 1;
    r.ma = offs[row + xlo];
    r.mb = offs[row + xhi + 1];
    if (per[0] && cx == 0) { r.wa = offs[row + nx - 1]; r.wb = offs[row + nx]; r.wsx = -R.L[0]; }
    else if (per[0] && cx == nx - 1) { r.wa = offs[row]; r.wb = offs[row + 1]; r.wsx = R.L[0]; }
    r.ok = true;
    return r;
}

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}
// LDS written by some lanes of a wave, then read by others
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}


// [7,9] density of moving particles (and their neighbour lists); velocity
// of Bound2 walls from the predicted velocities of their moving neighbours.
// A wave's 64 slots lie in a few consecutive cubes of one row, so for each
// of the 9 neighbour rows the union of its lanes' main ranges is one short
// slot interval: the wave loads it into LDS once (coalesced) and every lane
// scans its own sub-range there instead of issuing ~10 scattered global
// loads per row.  A union longer than TSTAGE, and the periodic image cubes,
// are read from global memory.  Scan order (image cube, then main range,
// row by row) is the same either way.  Rows without an image shift in any
// lane skip the shift adds (x + 0 is x: the same bits).
// The neighbour list of a wave is one block of 64 x NL_CAP entries, entry k
// of lane l at 64 k + l (coalesced rows for the force pass).
// LB (skin mode, a list rebuild): the same scan writes the Verlet list of
// every particle of a rebuilding replica (every pair within h + skin; for a
// Bound2 wall its moving partners only; VL_CAP entries per particle, blocks
// of VL_ROWS rows) and computes no density.
template <bool ENS, int TBS, int TSTAGE, bool LB = false>
__global__ __launch_bounds__(TBS) void k2_density(int n, GcRep R0, const GcRep* __restrict__ reps,
    const int* __restrict__ rep_of, const float4* __restrict__ posid, float4* __restrict__ velrho,
    const int* __restrict__ offs, const int* __restrict__ key, int* __restrict__ nl,
    int* __restrict__ nlc, int* __restrict__ nlover, const int* __restrict__ gate,
    const int* __restrict__ rbflag, tfloat3* __restrict__ rbpos, const unsigned char* __restrict__ gact)
{
    constexpr int CAP = LB ? VL_CAP : NL_CAP;
    constexpr int ROWS = LB ? VL_ROWS : NL_CAP;
    if (LB && *gate == 0) return;   // no rebuild this step
    __shared__ float4 stage[TBS / 64][TSTAGE > 0 ? TSTAGE : 1];
    const int lane = threadIdx.x & 63;
    float4* st = stage[threadIdx.x >> 6];
    // one group of TBS slots; a batch's list build strides over a fixed grid
    // (the rebuilding replicas' slots are contiguous runs, which a grid of
    // one group per TBS slots would place on one or two XCDs)
    auto group = [&](int lb) {
    const int p = lb * TBS + threadIdx.x;
    const bool inb = p < n;
    const float4 pi = inb ? posid[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int id = inb ? pid_of(pi.w) : 0;
    bool act = inb && id >= R0.P.nb2;
    if (LB && ENS && act) act = rbflag[rep_of[id]] != 0;   // lists of the other replicas stay
    if (__ballot(act) == 0) return;   // wave-uniform
    with_rep<ENS>(R0, reps, rep_of, inb, id, [&](const GcRep& R) __attribute__((always_inline)) {
    const GcParams& P = R.P;
    int cx = 0, cy = 0, cz = 0;
    if (act) key_to_cc(key[p], R, cx, cy, cz);
    const bool mov = act && (LB || id >= P.nf);
    const float ih = 1.0f / P.h;
    float d = 0.0f;  // self term added below
    float sw = 0.f, svx = 0.f, svy = 0.f, svz = 0.f;
    int* wl = nl + (size_t)__builtin_amdgcn_readfirstlane(p >> 6) * (ROWS * 64);
    unsigned woff = (unsigned)lane;   // = lane + 64 (entries so far)
    const unsigned wlast = (unsigned)((CAP - 1) * 64 + lane);   // overflow rewrites the last entry
    const float r2max = LB ? R.list_r2 : P.h2;
    auto hit = [&](int j, float r2, int idj, int code) {
        wl[min(woff, wlast)] = (j << 6) | code;
        woff += 64;
        float q = __builtin_amdgcn_sqrtf(r2) * ih, om = 1.0f - q;
        float w = (1.0f + 3.0f * q) * om * om * om;
        d += idj >= P.nb2 ? w : 0.0f;
    };
    auto consider = [&](auto SH, int j, float4 pj, float sx, float sy, float sz, int code) {
        float rx, ry, rz;
        if constexpr (decltype(SH)::value) {
            rx = pdiff(pi.x, pj.x, sx); ry = pdiff(pi.y, pj.y, sy); rz = pdiff(pi.z, pj.z, sz);
        } else {
            rx = pi.x - pj.x; ry = pi.y - pj.y; rz = pi.z - pj.z;
        }
        float r2 = rx * rx + ry * ry + rz * rz;
        const int idj = pid_of(pj.w);
        if (LB) {
            // a Bound2 wall lists its moving neighbours only (its velocity
            // is their weighted mean; wall-wall pairs enter no sum)
            if (mov && r2 < r2max && j != p && (id >= P.nf || idj >= P.nf)) {
                wl[min(woff, wlast)] = (j << 6) | code;
                woff += 64;
            }
        } else if (mov) {
            if (r2 < P.h2 && j != p) hit(j, r2, idj, code);
        } else if (r2 < P.h2 && idj >= P.nf) {
            float q = __builtin_amdgcn_sqrtf(r2) * ih, om = 1.0f - q;
            float w = (1.0f + 3.0f * q) * om * om * om;
            float4 vj = velrho[j];
            sw += w; svx += w * vj.x; svy += w * vj.y; svz += w * vj.z;
        }
    };
    for (int q = 0; q < 9; ++q) {
        RowRanges rr = act ? row_ranges(R, offs, cx, cy, cz, q) : RowRanges{0, 0, 0, 0, 0.f, 0.f, 0.f, false};
        const int lo = wave_min(rr.ok && rr.mb > rr.ma ? rr.ma : 0x7fffffff);
        const int hi = wave_max(rr.ok && rr.mb > rr.ma ? rr.mb : -1);
        if (rr.ok && rr.wb > rr.wa) {
            const int code = shift_code(rr.wsx, rr.sy, rr.sz);
            scan_range(posid, rr.wa, rr.wb, [&](int j, float4 pj) {
                consider(std::true_type{}, j, pj, rr.wsx, rr.sy, rr.sz, code);
            });
        }
        const int code = shift_code(0.0f, rr.sy, rr.sz);
        auto main_range = [&](auto SH) {
            if (TSTAGE > 0 && hi > lo && hi - lo <= TSTAGE) {   // wave-uniform
                for (int i = lane; i < hi - lo; i += 64) st[i] = posid[lo + i];
                wave_lds_sync();
                if (rr.ok) {
                    int j = rr.ma;
                    for (; j + 4 <= rr.mb; j += 4) {
                        float4 a = st[j - lo], b = st[j + 1 - lo], c = st[j + 2 - lo], e4 = st[j + 3 - lo];
                        consider(SH, j, a, 0.0f, rr.sy, rr.sz, code);
                        consider(SH, j + 1, b, 0.0f, rr.sy, rr.sz, code);
                        consider(SH, j + 2, c, 0.0f, rr.sy, rr.sz, code);
                        consider(SH, j + 3, e4, 0.0f, rr.sy, rr.sz, code);
                    }
                    for (; j < rr.mb; ++j) consider(SH, j, st[j - lo], 0.0f, rr.sy, rr.sz, code);
                }
                wave_lds_sync();
            } else if (rr.ok) {
                scan_range(posid, rr.ma, rr.mb, [&](int j, float4 pj) {
                    consider(SH, j, pj, 0.0f, rr.sy, rr.sz, code);
                });
            }
        };
        if (__ballot(rr.ok && (rr.sy != 0.0f || rr.sz != 0.0f))) main_range(std::true_type{});
        else main_range(std::false_type{});
    }
    if (LB) {
        if (mov) {
            const int cnt = (int)((woff - (unsigned)lane) >> 6);
            nlc[p] = cnt <= CAP ? cnt : -1;
            if (cnt > CAP) atomicAdd(nlover, 1);
            rbpos[p] = tfloat3{pi.x, pi.y, pi.z};   // displacements are measured from here
        }
    } else if (mov) {
        const int cnt = (int)((woff - (unsigned)lane) >> 6);
        velrho[p].w = P.m * P.wnorm * (d + 1.0f);
        nlc[p] = cnt <= NL_CAP ? cnt : -1;
        if (cnt > NL_CAP) atomicAdd(nlover, 1);
    } else if (act && id < P.nf) {
        const float3 u = wall_velocity(P, pi.y, sw, svx, svy, svz);
        float4 v = velrho[p];
        v.x = u.x; v.y = u.y; v.z = u.z;
        velrho[p] = v;
    }
    });
    };
    if (LB && ENS) for_groups_of<TBS>((n + TBS - 1) / TBS, gact, group);
    else group(xcd_block((n + TBS - 1) / TBS));
}

// [10-13,17] SDPD pair force with the pressure from the EOS on the fly,
// Bound1 repulsion, body force; corrector for fluid (membrane vertices get
// it after their membrane force).
// VL (skin mode): the neighbours are the entries of the particle's Verlet
// list whose bits the density pass set (r < h at this step), in list order.
template <bool ENS, int TBS, bool VL = false>
__global__ __launch_bounds__(TBS) void k2_force(int nmov, GcRep R0, const GcRep* __restrict__ reps,
    const int* __restrict__ rep_of, int driveon, const int* __restrict__ act, const float4* __restrict__ posid,
    const float4* __restrict__ velrho, const int* __restrict__ offs, const int* __restrict__ key,
    const int* __restrict__ nl, const int* __restrict__ nlc,
    const tfloat3* __restrict__ velhalf, tfloat3* __restrict__ force, float4* __restrict__ vout,
    const unsigned long long* __restrict__ ctr, int* __restrict__ fault)
{
    // one lane per moving particle (walls get no force: no lane idles on a
    // wall slot); act lists their slots in slot order
    const int t = xcd_block((nmov + TBS - 1) / TBS) * TBS + threadIdx.x;
    if (t >= nmov) return;
    const int p = act[t];
    float4 pi = posid[p];
    int id = pid_of(pi.w);
    const int oi = VL ? woff_of(pi.w) : 0;
    if (id < R0.P.nf) return;   // a wall slot (only with moving walls): no force
    with_rep<ENS>(R0, reps, rep_of, true, id, [&](const GcRep& R) __attribute__((always_inline)) {
    const GcParams& P = R.P;   // a reference: a copy of the batch record would live in scratch
    const unsigned long long c = *ctr - 1ull;   // advanced by this step's offsets pass
    const uint32_t ctr_lo = (uint32_t)c, ctr_hi = (uint32_t)(c >> 32);
    const uint32_t seed_lo = P.seed_lo, seed_hi = P.seed_hi;
    const int lid = ENS ? local_id(R, id) : id;
    // the replica's id offsets and periods in registers (one table read)
    const int lo0 = R.id_off[0] - R.loc_start[0], lo1 = R.id_off[1] - R.loc_start[1];
    const int lo2 = R.id_off[2] - R.loc_start[2], lo3 = R.id_off[3] - R.loc_start[3];
    const float L0 = R.L[0], L1 = R.L[1], L2 = R.L[2];
    float4 vi = velrho[p];
    float di = vi.w * P.inv_m;
    // P_i / d_i^2 with P = c^2 rho and d = rho / m is c^2 m / d_i; the
    // reciprocals are v_rcp_f32 / v_rsq_f32 (1 ulp), well inside the 1e-5
    // force tolerance and much cheaper than IEEE division
    const float rdi = __builtin_amdgcn_rcpf(di);
    float pi_d2 = (id < P.nc ? P.c1sq : P.c2sq) * P.m * rdi;
    int cx, cy, cz;
    key_to_cc(key[p], R, cx, cy, cz);
    const float ih = 1.0f / P.h;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    auto pair = [&](int code, const float4 pj, const float4 vj) {
        const int oj = VL ? woff_of(pj.w) : 0;
        float rx, ry, rz;
        if (__ballot((code & 63) | oj | oi)) {   // an image neighbour in the wave (x + 0 is x: same bits)
            float sx, sy, sz;
            if (VL) {
                sx = eff_shift(code, oi, oj, 0, L0); sy = eff_shift(code, oi, oj, 1, L1);
                sz = eff_shift(code, oi, oj, 2, L2);
            } else {
                sx = code_shift(code, L0); sy = code_shift(code >> 2, L1); sz = code_shift(code >> 4, L2);
            }
            rx = pdiff(pi.x, pj.x, sx); ry = pdiff(pi.y, pj.y, sy); rz = pdiff(pi.z, pj.z, sz);
        } else {
            rx = pi.x - pj.x; ry = pi.y - pj.y; rz = pi.z - pj.z;
        }
        float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 == 0.f) return;  // coincident particles exert no pair force
        int idj = pid_of(pj.w);
        float ir = __builtin_amdgcn_rsqf(r2);
        float r = r2 * ir;
        float om = 1.0f - r * ih;
        float F = P.fnorm * om * om;
        if (idj < P.nb2) {  // Bound1 wall: repulsion only
            float kk = P.krep * F;
            fx += kk * rx; fy += kk * ry; fz += kk * rz;
            return;
        }
        float dj = vj.w * P.inv_m;
        float rdj = __builtin_amdgcn_rcpf(dj);
        float cC = (pi_d2 + (idj < P.nc ? P.c1sq : P.c2sq) * P.m * rdj) * F;
        float ex = rx * ir, ey = ry * ir, ez = rz * ir;
        float vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
        float ev = ex * vx + ey * vy + ez * vz;
        float gg = F * rdi * rdj;
        float ga = gg * P.visc_a, gb = gg * P.visc_b * ev;
        fx += cC * rx - (ga * vx + gb * ex);
        fy += cC * ry - (ga * vy + gb * ey);
        fz += cC * rz - (ga * vz + gb * ez);
        if (P.noise_on) {
            int ljd = ENS ? local_id_of(idj, P.nb2, P.nf, P.nc, lo0, lo1, lo2, lo3) : idj;
            float n0, n1, n2;
            pair_normals_k(seed_lo, seed_hi, ctr_lo, ctr_hi, (uint32_t)lid, (uint32_t)ljd, n0, n1, n2);
            float exi = ex * n0 + ey * n1 + ez * n2;
            float amp = __builtin_amdgcn_sqrtf(P.two_kT_over_dt * gg);
            amp = lid < ljd ? amp : -amp;
            float pa = amp * P.sq_a, pe = amp * (P.sq_ab - P.sq_a) * exi;
            fx += pa * n0 + pe * ex;
            fy += pa * n1 + pe * ey;
            fz += pa * n2 + pe * ez;
        }
    };
    // The density pass listed the neighbours (r < h, j != p) in the scan
    // order of the cubes, so the sums below run in that order either way.
    const int cnt = nlc[p];
    if (cnt >= 0 && VL) {
        const int* wl = nl + (size_t)(p >> 6) * (VL_ROWS * 64) + (p & 63);   // the list's block
        const int nw = (cnt + 31) >> 5;
        int w = 0;
        unsigned m = nw > 0 ? (unsigned)wl[64 * VL_CAP] : 0u;
        // the entry index of the next set bit, -1 after the last
        auto next_k = [&]() -> int {
            while (m == 0u) {
                if (++w >= nw) return -1;
                m = (unsigned)wl[64 * (VL_CAP + w)];
            }
            const int b = __builtin_ctz(m);
            m &= m - 1u;
            return 32 * w + b;
        };
        // the next pair's entry, position and velocity are loaded during the
        // current pair (10M: 1,248 -> 1,205 us; also loading the entry after
        // next was slower, 1,232)
        int k = next_k();
        int code = k >= 0 ? wl[64 * k] : 0;
        float4 pjn = posid[code >> 6], vjn = velrho[code >> 6];
        while (k >= 0) {
            const int cur = code;
            const float4 pj = pjn, vj = vjn;
            k = next_k();
            if (k >= 0) {
                code = wl[64 * k];
                pjn = posid[code >> 6];
                vjn = velrho[code >> 6];
            }
            pair(cur, pj, vj);
        }
    } else if (cnt >= 0) {
        const int* wl = nl + (size_t)(p >> 6) * (NL_CAP * 64) + (p & 63);   // the density lane's block
        int next = cnt > 0 ? wl[0] : 0;   // the next entry is loaded during the current pair
        for (int k = 0; k < cnt; ++k) {
            const int code = next;
            if (k + 1 < cnt) next = wl[64 * (k + 1)];
            pair(code, posid[code >> 6], velrho[code >> 6]);
        }
    } else {  // more than the list holds: scan the cubes again
        visit2(R, offs, cx, cy, cz, [&](int jb, int je, float sx, float sy, float sz) {
            const int code = shift_code(sx, sy, sz);
            for (int j = jb; j < je; ++j) {
                float4 pj = posid[j];
                if (VL) {   // the cubes of the last rebuild, wraps since added
                    const int oj = woff_of(pj.w);
                    sx = eff_shift(code, oi, oj, 0, L0); sy = eff_shift(code, oi, oj, 1, L1);
                    sz = eff_shift(code, oi, oj, 2, L2);
                }
                float rx = pdiff(pi.x, pj.x, sx), ry = pdiff(pi.y, pj.y, sy), rz = pdiff(pi.z, pj.z, sz);
                float r2 = rx * rx + ry * ry + rz * rz;
                if (r2 < P.h2 && j != p) pair((j << 6) | code, pj, velrho[j]);
            }
        });
    }
    if (driveon) fx += ext_force_x(P, pi.y);
    if (!isfinite(fx + fy + fz)) atomicOr(fault, GC_FAULT_NONFINITE);
    force[p] = tfloat3{fx, fy, fz};
    if (id < P.nc) {
        float h = 0.5f * P.dt * P.inv_m;
        tfloat3 vh = velhalf[p];
        vout[p] = make_float4(vh.x + h * fx, vh.y + h * fy, vh.z + h * fz, vi.w);
    }
    });
}

// [16] Morse between membrane vertices of different cells
// Iterates the density pass's neighbour list, or in skin mode the Verlet
// list (MaxCut <= CutRadius, so every Morse partner is in it, in scan
// order); a vertex whose list overflowed scans its cubes.
// One thread per membrane vertex (its slot from CurPos).  SK: skin mode
// (Verlet lists, wrap codes).
template <bool ENS, bool SK>
__global__ __launch_bounds__(BS) void k2_agg(int nm, GcRep R0, const GcRep* __restrict__ reps,
    const int* __restrict__ rep_of, const int* __restrict__ curpos, const float4* __restrict__ posid,
    const int* __restrict__ offs, const int* __restrict__ key, const int* __restrict__ nl,
    const int* __restrict__ nlc, const int* __restrict__ celltab, tfloat3* __restrict__ force,
    tfloat3* __restrict__ aggforce, int nl_stride)
{
    const int v = xcd_block((nm + BS - 1) / BS) * BS + threadIdx.x;
    if (v >= nm) return;
    const int p = curpos[R0.P.nc + v];
    float4 pi = posid[p];
    int id = pid_of(pi.w);
    const int oi = SK ? woff_of(pi.w) : 0;
    with_rep<ENS>(R0, reps, rep_of, true, id, [&](const GcRep& R) __attribute__((always_inline)) {
    const GcParams& P = R.P;
    if (P.De == 0.0f) return;
    int cell = celltab[id - P.nc];
    int cx, cy, cz;
    key_to_cc(key[p], R, cx, cy, cz);
    const float tb = 2.0f * P.beta * P.De;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    // the image shift of the pair (list code or scanned cube), wraps since
    // the last rebuild added in skin mode
    auto shifts = [&](int code, float4 pj, float& sx, float& sy, float& sz) {
        if (SK) {
            const int oj = woff_of(pj.w);
            sx = eff_shift(code, oi, oj, 0, R.L[0]); sy = eff_shift(code, oi, oj, 1, R.L[1]);
            sz = eff_shift(code, oi, oj, 2, R.L[2]);
        } else {
            sx = code_shift(code, R.L[0]); sy = code_shift(code >> 2, R.L[1]); sz = code_shift(code >> 4, R.L[2]);
        }
    };
    auto morse = [&](int j, int code) {
        float4 pj = posid[j];
        int idj = pid_of(pj.w);
        if (idj < P.nc) return;
        float sx, sy, sz;
        shifts(code, pj, sx, sy, sz);
        float rx = pdiff(pi.x, pj.x, sx), ry = pdiff(pi.y, pj.y, sy), rz = pdiff(pi.z, pj.z, sz);
        float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 >= P.mcut2 || r2 == 0.f || celltab[idj - P.nc] == cell) return;
        float ir = __builtin_amdgcn_rsqf(r2);
        float r = r2 * ir;
        float ex = expf(P.beta * (P.r0 - r));
        float fm = tb * (ex * ex - ex) * ir;
        fx += fm * rx; fy += fm * ry; fz += fm * rz;
    };
    const int cnt = nlc[p];
    if (cnt >= 0) {
        // the lanes' slots are scattered: each lane's own list block.  Eight
        // entries at a time: their codes, then their positions, then their
        // cells are loaded from clamped indices before any is used (one entry
        // after another, each pair would wait out three dependent loads);
        // the pairs are then summed in list order
        const int* wl = nl + (size_t)(p >> 6) * nl_stride + (p & 63);
        constexpr int G = 8;
        for (int k0 = 0; k0 < cnt; k0 += G) {
            int code[G], cj[G];
            float4 pj[G];
#pragma unroll
            for (int u = 0; u < G; ++u) code[u] = wl[64 * min(k0 + u, cnt - 1)];
#pragma unroll
            for (int u = 0; u < G; ++u) pj[u] = posid[code[u] >> 6];
#pragma unroll
            for (int u = 0; u < G; ++u) cj[u] = celltab[max(pid_of(pj[u].w) - P.nc, 0)];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                if (k0 + u >= cnt) break;
                const int idj = pid_of(pj[u].w);
                if (idj < P.nc) continue;
                float sx, sy, sz;
                shifts(code[u], pj[u], sx, sy, sz);
                const float rx = pdiff(pi.x, pj[u].x, sx), ry = pdiff(pi.y, pj[u].y, sy),
                            rz = pdiff(pi.z, pj[u].z, sz);
                const float r2 = rx * rx + ry * ry + rz * rz;
                if (r2 >= P.mcut2 || r2 == 0.f || cj[u] == cell) continue;
                const float ir = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * ir;
                const float ex = expf(P.beta * (P.r0 - r));
                const float fm = tb * (ex * ex - ex) * ir;
                fx += fm * rx; fy += fm * ry; fz += fm * rz;
            }
        }
    } else {
        visit2(R, offs, cx, cy, cz, [&](int jb, int je, float sx, float sy, float sz) {
            const int code = shift_code(sx, sy, sz);
            for (int j = jb; j < je; ++j) morse(j, code);
        });
    }
    aggforce[id - P.nc] = tfloat3{fx, fy, fz};
    tfloat3 F = force[p];
    F.x += fx; F.y += fy; F.z += fz;
    force[p] = F;
    });
}

// ---------------------------------------------------------------------------
// Skin mode (Verlet lists).  Between rebuilds the slots, cube keys, CurPos
// and the moving-slot map stay as the last rebuild left them, and the lists
// hold every pair within h + skin at the rebuild.  Positions are advanced
// and wrapped in place every step; each particle carries its wraps since the
// rebuild (the wrap code in the id word of posid), so a list entry's image
// shift S becomes S + (o_j - o_i) L, the shift of the wrapped pair.  A pair
// not in the lists was more than h + skin apart, so it cannot be within h
// while its relative displacement since the rebuild is at most skin: the
// rebuild test bounds that by every |D_i| <= skin / 2 or by the bounding box
// of all D_i (a common translation moves no pair).

__device__ __forceinline__ float wave_maxf(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Displacement statistics of a workgroup's particles since the last
// rebuild, DSTAT non-negative values: the largest squared displacement
// |D|^2 and, per axis, the largest max(D, 0) and max(-D, 0) (the bounding
// box of the displacements, which holds D = 0 too).  Into bmax[DSTAT b ..]
// (and the group's replica into brep[b], -1 when its particles belong to
// several: those go to their replicas' words of disp2 by atomics, float
// bits of a non-negative value ordering as unsigned integers).  Called by
// every thread of the group; WAVES = waves per group.
constexpr int DSTAT = DSTAT_N;
__device__ __forceinline__ void disp_stats(float dx, float dy, float dz, float st[DSTAT])
{
    float dd = dx * dx + dy * dy + dz * dz;
    st[0] = dd <= 3.0e38f ? dd : 3.0e38f;   // non-finite: forces a rebuild (which flags it)
    st[1] = fmaxf(dx, 0.0f); st[2] = fmaxf(-dx, 0.0f);
    st[3] = fmaxf(dy, 0.0f); st[4] = fmaxf(-dy, 0.0f);
    st[5] = fmaxf(dz, 0.0f); st[6] = fmaxf(-dz, 0.0f);
    if (!(st[0] < 3.0e38f))
        for (int k = 1; k < DSTAT; ++k) st[k] = 3.0e38f;
}

template <bool ENS, int WAVES>
__device__ __forceinline__ void block_disp(bool live, const float st[DSTAT], int r, int b,
                                           float* __restrict__ bmax, int* __restrict__ brep,
                                           unsigned* __restrict__ disp2)
{
    __shared__ float smax[WAVES][DSTAT];
    __shared__ int srep[WAVES];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int r0 = 0;
    bool mixed = false;
    if constexpr (ENS) {
        r0 = __builtin_amdgcn_readfirstlane(live ? r : -1);
        if (r0 < 0) {   // the first lane is not live: take the first live lane's replica
            const unsigned long long lv = __ballot(live);
            r0 = lv ? __shfl(r, (int)__ffsll((long long)lv) - 1) : -1;
        }
        mixed = __ballot(live && r != r0) != 0;
        if (mixed && live)
            for (int k = 0; k < DSTAT; ++k)
                if (st[k] > 0.0f) atomicMax(&disp2[DSTAT * r + k], __float_as_uint(st[k]));
    }
#pragma unroll
    for (int k = 0; k < DSTAT; ++k) {
        const float m = wave_maxf(live && !mixed ? st[k] : 0.0f);
        if (lane == 0) smax[w][k] = m;
    }
    if (lane == 0) srep[w] = mixed ? -1 : r0;
    __syncthreads();
    if (threadIdx.x == 0) {
        float bm[DSTAT];
        for (int k = 0; k < DSTAT; ++k) bm[k] = 0.0f;
        int rep = -2;   // -2: no particle yet
        bool split = false;
        for (int q = 0; q < WAVES; ++q) {
            for (int k = 0; k < DSTAT; ++k) bm[k] = fmaxf(bm[k], smax[q][k]);
            if (ENS && srep[q] >= 0) {
                if (rep == -2) rep = srep[q];
                else if (rep != srep[q]) split = true;
            }
        }
        if (ENS && split) {   // several replicas in the group (rare): their own words
            for (int q = 0; q < WAVES; ++q)
                if (srep[q] >= 0)
                    for (int k = 0; k < DSTAT; ++k)
                        if (smax[q][k] > 0.0f) atomicMax(&disp2[DSTAT * srep[q] + k], __float_as_uint(smax[q][k]));
            for (int k = 0; k < DSTAT; ++k) bm[k] = 0.0f;
            rep = -1;
        }
        for (int k = 0; k < DSTAT; ++k) bmax[DSTAT * b + k] = bm[k];
        if (ENS) brep[b] = rep < 0 ? -1 : rep;
    }
}

// [1,2,5] skin mode, every step: r += dt v + dt^2/(2m) F in place (moving
// particles and moving walls) with the periodic wrap and image counters of
// the per-step gather (the same expressions: the same bits), the
// Groot-Warren predictor (velhalf, velrho), the FP32 neighbour-loop copy of
// the position (posid, with the particle's wraps since the last rebuild in
// its wrap code; coor for membrane vertices, which the membrane kernels
// read), and the largest squared displacement since the last rebuild (from
// the position saved then) per workgroup.  Frozen walls change nothing.
// Membrane vertices under substepping were moved and wrapped by the
// previous step's substeps (which noted their displacement and wraps): they
// take their velocity as is.
template <bool ENS, int TBS>
__global__ __launch_bounds__(TBS) void k3_advance(int n, GcRep R0, const GcRep* __restrict__ reps,
    const int* __restrict__ rep_of, int mmove, int wmove, double* __restrict__ pos64,
    const float4* __restrict__ vout, const tfloat3* __restrict__ force, const int* __restrict__ sp,
    tint3* __restrict__ nt, unsigned char* __restrict__ woff, const tfloat3* __restrict__ rbpos,
    float4* __restrict__ posid, tfloat3* __restrict__ coor, float4* __restrict__ velrho,
    tfloat3* __restrict__ velhalf, float* __restrict__ bmax, int* __restrict__ brep,
    unsigned* __restrict__ disp2)
{
    const int nb = (n + TBS - 1) / TBS;
    const int b = xcd_block(nb);
    if (b >= nb) return;   // a padding group of the XCD mapping (whole group)
    const int p = b * TBS + threadIdx.x;
    const bool inb = p < n;
    const int id = inb ? sp[p] : 0;
    float st[DSTAT] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    with_rep<ENS>(R0, reps, rep_of, inb, id, [&](const GcRep& R) __attribute__((always_inline)) {
    if (!inb) return;
    const GcParams& P = R.P;
    const bool mv = moves(P, id, mmove), wmv = wall_moves(P, id, wmove);
    if (!mv && !wmv && id < P.nf) return;   // a frozen wall
    double u[3] = {pos64[p], pos64[n + p], pos64[2 * n + p]};
    float4 v = vout[p];
    int wc = woff[p];
    if (mv || wmv) {
        if (mv) {
            const tfloat3 f = force[p];
            advance_step(P, v, f, u);
            float h = 0.5f * P.dt * P.inv_m, l = P.lam * P.dt * P.inv_m;
            velhalf[p] = tfloat3{v.x + h * f.x, v.y + h * f.y, v.z + h * f.z};
            v.x += l * f.x; v.y += l * f.y; v.z += l * f.z;
            velrho[p] = v;
        } else {
            advance_step(P, wall_v4(P, u[1]), tfloat3{0.f, 0.f, 0.f}, u);
        }
        int img[3] = {0, 0, 0};
        wrap_once(R, P, u, img);
        pos64[p] = u[0]; pos64[n + p] = u[1]; pos64[2 * n + p] = u[2];
        if (img[0] | img[1] | img[2]) {
            tint3 t = nt[p];
            t.x += img[0]; t.y += img[1]; t.z += img[2];
            nt[p] = t;
            wc = coef_code(code_coef(wc, 0) + img[0], code_coef(wc, 1) + img[1], code_coef(wc, 2) + img[2]);
            woff[p] = (unsigned char)wc;
        }
        // displacement since the rebuild, in the rebuild's image frame
        const tfloat3 r0 = rbpos[p];
        const float dx = ((float)u[0] + (float)code_coef(wc, 0) * R.L[0]) - r0.x;
        const float dy = ((float)u[1] + (float)code_coef(wc, 1) * R.L[1]) - r0.y;
        const float dz = ((float)u[2] + (float)code_coef(wc, 2) * R.L[2]) - r0.z;
        disp_stats(dx, dy, dz, st);
    } else {
        velhalf[p] = tfloat3{v.x, v.y, v.z};
        velrho[p] = v;
    }
    const float fx = (float)u[0], fy = (float)u[1], fz = (float)u[2];
    posid[p] = make_float4(fx, fy, fz, id_word(id, wc));
    if (id >= P.nc) coor[p] = tfloat3{fx, fy, fz};
    });
    block_disp<ENS, TBS / 64>(inb, st, ENS && inb ? rep_of[id] : 0, b, bmax, brep, disp2);
}

// Skin mode, once per step after the advance: the workgroup displacement
// statistics (the advance's groups, then the previous step's membrane
// substeps' from ms_off on) reduced per replica into disp2 (which also holds
// the atomics of groups spanning replicas).  Each thread takes a contiguous
// run of entries and flushes its maxima when the replica changes.
constexpr int REDUCE_BS = 256, REDUCE_NB = 64;
__global__ __launch_bounds__(REDUCE_BS) void k3_reduce(int nrep, const float* __restrict__ bmax,
    const int* __restrict__ brep, int nb_adv, int ms_off, int nb_ms, unsigned* __restrict__ disp2)
{
    const int E = nb_adv + nb_ms, T = REDUCE_NB * REDUCE_BS;
    const int C = (E + T - 1) / T;
    const int t = blockIdx.x * REDUCE_BS + threadIdx.x;
    float m[DSTAT];
    for (int k = 0; k < DSTAT; ++k) m[k] = 0.0f;
    int cur = -1;
    auto flush = [&]() {
        if (cur >= 0)
            for (int k = 0; k < DSTAT; ++k)
                if (m[k] > 0.0f) atomicMax(&disp2[DSTAT * cur + k], __float_as_uint(m[k]));
        for (int k = 0; k < DSTAT; ++k) m[k] = 0.0f;
    };
    for (int e = t * C; e < min(E, (t + 1) * C); ++e) {
        // entries past the advance's groups are the substeps' (one suspension)
        const int q = e < nb_adv ? e : ms_off + (e - nb_adv);
        const int r = nrep == 1 ? 0 : (e < nb_adv ? brep[q] : 0);
        if (r < 0) continue;   // a group spanning replicas: already in disp2
        if (r != cur) { flush(); cur = r; }
        for (int k = 0; k < DSTAT; ++k) m[k] = fmaxf(m[k], bmax[DSTAT * q + k]);
    }
    if (nrep == 1) {   // one word set: reduce the group first
        __shared__ float sm[REDUCE_BS / 64][DSTAT];
        for (int k = 0; k < DSTAT; ++k) {
            const float v = wave_maxf(m[k]);
            if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6][k] = v;
        }
        __syncthreads();
        if (threadIdx.x < DSTAT) {
            float v = 0.0f;
            for (int w = 0; w < REDUCE_BS / 64; ++w) v = fmaxf(v, sm[w][threadIdx.x]);
            if (v > 0.0f) atomicMax(&disp2[threadIdx.x], __float_as_uint(v));
        }
    } else {
        // a wave whose lanes end on one replica (nearly all) reduces first:
        // one atomic per statistic instead of 64 on the same word
        const int c0 = __builtin_amdgcn_readfirstlane(cur);
        if (__ballot(cur >= 0 && cur != c0) == 0) {
            for (int k = 0; k < DSTAT; ++k) {
                const float v = wave_maxf(cur >= 0 ? m[k] : 0.0f);
                if ((threadIdx.x & 63) == 0 && c0 >= 0 && v > 0.0f)
                    atomicMax(&disp2[DSTAT * c0 + k], __float_as_uint(v));
            }
        } else {
            flush();
        }
    }
}

// Skin mode, a batch: workgroup r reduces the statistics of replica r's
// advance groups (its slots are contiguous: groups slotbase[r] / gsz to
// (slotbase[r + 1] - 1) / gsz, gsz the advance's group size); the groups
// spanning two replicas put theirs into disp2 directly.  One atomic per
// statistic and replica.
__global__ __launch_bounds__(REDUCE_BS) void k3_reduce_rep(const float* __restrict__ bmax,
    const int* __restrict__ brep, const int* __restrict__ slotbase, int gsz, unsigned* __restrict__ disp2)
{
    const int r = blockIdx.x;
    const int g0 = slotbase[r] / gsz, g1 = (slotbase[r + 1] - 1) / gsz;
    float m[DSTAT];
    for (int k = 0; k < DSTAT; ++k) m[k] = 0.0f;
    for (int q = g0 + (int)threadIdx.x; q <= g1; q += REDUCE_BS) {
        if (brep[q] != r) continue;
        for (int k = 0; k < DSTAT; ++k) m[k] = fmaxf(m[k], bmax[DSTAT * q + k]);
    }
    __shared__ float sm[REDUCE_BS / 64][DSTAT];
    for (int k = 0; k < DSTAT; ++k) {
        const float v = wave_maxf(m[k]);
        if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < DSTAT) {
        float v = 0.0f;
        for (int w = 0; w < REDUCE_BS / 64; ++w) v = fmaxf(v, sm[w][threadIdx.x]);
        if (v > 0.0f) atomicMax(&disp2[DSTAT * r + threadIdx.x], __float_as_uint(v));
    }
}

// Skin mode, per replica from its displacement statistics since its last
// rebuild: the lists stay valid while every pair's relative displacement
// |D_i - D_j| is at most the skin, which holds if every |D_i| <= skin/2 or
// if the bounding box of all D (a uniform translation moves no pair) has a
// diagonal <= skin; when both fail (or a value is non-finite) the replica's
// lists are rebuilt this step.  Clears the statistics and advances the
// pair-noise counter (the force pass reads the value before the advance).
__global__ __launch_bounds__(64) void k3_decide(int nrep, const GcRep* __restrict__ reps,
    unsigned* __restrict__ disp2, int* __restrict__ rbflag, int* __restrict__ rbany, int* __restrict__ nrebuild,
    unsigned long long* __restrict__ ctr)
{
    int any = 0;
    for (int r = threadIdx.x; r < nrep; r += 64) {
        float v[DSTAT];
        for (int k = 0; k < DSTAT; ++k) {
            v[k] = __uint_as_float(disp2[DSTAT * r + k]);
            disp2[DSTAT * r + k] = 0u;
        }
        const float hs = reps[r].half_skin;
        const float bx = v[1] + v[2], by = v[3] + v[4], bz = v[5] + v[6];
        const bool each = sqrtf(v[0]) <= hs;                          // every |D_i| <= skin/2
        const bool box = sqrtf(bx * bx + by * by + bz * bz) <= 2.0f * hs;   // diagonal <= skin
        const int rb = !(each || box);
        rbflag[r] = rb;
        if (rb) { nrebuild[r] += 1; any = 1; }
    }
    any = __ballot(any) != 0;
    if (threadIdx.x == 0) { *rbany = any; *ctr += 1ull; }
}

// [3] skin mode, a rebuild: cube key of each particle of a rebuilding
// replica from its position, histogram rank as in the per-step key pass.
// In a batch the other replicas keep their slots and their cube counts
// (which persist between rebuilds) and are skipped.
template <bool ENS, int TBS>
__global__ __launch_bounds__(TBS) void k3_keys(int n, GcRep R0, const GcRep* __restrict__ reps,
    const int* __restrict__ rep_of, const int* __restrict__ gate, const int* __restrict__ rbflag,
    const unsigned char* __restrict__ gact, const double* __restrict__ pos64, const int* __restrict__ sp,
    int* __restrict__ key, int* __restrict__ rank, int* __restrict__ cnt)
{
    if (*gate == 0) return;
    for_groups_of<TBS>((n + TBS - 1) / TBS, ENS ? gact : nullptr, [&](int lgb) {
    const int p = lgb * TBS + threadIdx.x;
    bool act = p < n;
    const int id = act ? sp[p] : 0;
    if (ENS && act) act = rbflag[rep_of[id]] != 0;
    if (__ballot(act) == 0) return;
    int k = -1;
    with_rep<ENS>(R0, reps, rep_of, act, id, [&](const GcRep& R) __attribute__((always_inline)) {
    if (!act) return;
    const GcParams& P = R.P;
    const double u[3] = {pos64[p], pos64[n + p], pos64[2 * n + p]};
    int cx = cube_coord((float)u[0], R.bbox[0], P.off[0], R.invw[0], R.nc[0]);
    int cy = cube_coord((float)u[1], R.bbox[2], P.off[1], R.invw[1], R.nc[1]);
    int cz = cube_coord((float)u[2], R.bbox[4], P.off[2], R.invw[2], R.nc[2]);
    k = R.cube_off + cx + R.nc[0] * cube_row(R, cy, cz);
    key[p] = k;
    });
    hist_rank(act, k, p, rank, cnt);
    });
}

// Skin mode, a batch rebuild: the rebuilding replicas' cube counts cleared
// before their key pass (the others' stay as their last rebuild left them).
__global__ __launch_bounds__(BS) void k3_clear_counts(int K, const int* __restrict__ gate,
    const int* __restrict__ rbflag, const int* __restrict__ cubebase, int nrep,
    const unsigned char* __restrict__ cact, int* __restrict__ cnt)
{
    if (*gate == 0) return;
    for_groups_of<BS>((K + BS - 1) / BS, cact, [&](int lb) {
    const int k = lb * BS + threadIdx.x;
    if (k < K && rbflag[base_rep(cubebase, nrep, k)]) cnt[k] = 0;
    });
}

// Skin mode, a batch rebuild: the activity bytes of the 64-slot chunks
// (entries [0, nsc)) and of the 64-cube chunks (then ncc) from the replicas
// each chunk spans (chunkrep: first and last) and their rebuild flags.
__global__ __launch_bounds__(BS) void k3_mark(int nsc, int ncc, const int* __restrict__ gate,
    const int* __restrict__ rbflag, const int2* __restrict__ chunkrep, unsigned char* __restrict__ sact,
    unsigned char* __restrict__ cact)
{
    if (*gate == 0) return;
    for_groups((nsc + ncc + BS - 1) / BS, [&](int lb) {
    const int c = lb * BS + threadIdx.x;
    if (c >= nsc + ncc) return;
    const int2 rr = chunkrep[c];
    int a = 0;
    for (int r = rr.x; r <= rr.y; ++r) a |= rbflag[r];
    if (c < nsc) sact[c] = a != 0;
    else cact[c - nsc] = a != 0;
    });
}

// Skin mode, a batch, after a host-driven build (which clears the counts):
// the cube counts restored from the offsets.
__global__ __launch_bounds__(BS) void k3_counts(int K, const int* __restrict__ offs, int* __restrict__ cnt)
{
    const int k = blockIdx.x * BS + threadIdx.x;
    if (k < K) cnt[k] = offs[k + 1] - offs[k];
}

// Skin mode, a rebuild: the gather wrote buffer 1; buffer 0 stays current.
__global__ __launch_bounds__(BS) void k3_copyback(int n, const int* __restrict__ gate,
    const double* __restrict__ pos1, double* __restrict__ pos0, const tint3* __restrict__ nt1,
    tint3* __restrict__ nt0, const int* __restrict__ sp1, int* __restrict__ sp0,
    const int* __restrict__ rep_of, const int* __restrict__ rbflag, const unsigned char* __restrict__ gact)
{
    if (*gate == 0) return;
    for_groups_of<BS>((n + BS - 1) / BS, gact, [&](int lb) {
    const int p = lb * BS + threadIdx.x;
    if (p >= n) return;
    if (rbflag && !rbflag[rep_of[sp0[p]]]) return;   // a batch: the rebuilt replicas' slots only
    pos0[p] = pos1[p]; pos0[n + p] = pos1[n + p]; pos0[2 * n + p] = pos1[2 * n + p];
    nt0[p] = nt1[p];
    sp0[p] = sp1[p];
    });
}

// [7,9] skin mode, every step: number density of moving particles and
// velocity of Bound2 walls over their Verlet lists.  The wave runs through
// the entries in lockstep (entry k of every lane is one coalesced row),
// eight entries' codes and then their positions loaded before use; r < h
// selects this step's neighbours, whose bits go into one row per 32 entries
// after the entries (read by the force pass).  The sums run in list order.
// A particle whose list overflowed scans its cubes (the rebuild's cubes
// are at least h + skin wide, so its neighbours are in the 27 around it).
template <bool ENS, int TBS>
__global__ __launch_bounds__(TBS) void k3_density(int n, GcRep R0, const GcRep* __restrict__ reps,
    const int* __restrict__ rep_of, const float4* __restrict__ posid, float4* __restrict__ velrho,
    const int* __restrict__ offs, const int* __restrict__ key, int* __restrict__ nl,
    const int* __restrict__ nlc)
{
    const int lane = threadIdx.x & 63;
    const int p = xcd_block((n + TBS - 1) / TBS) * TBS + threadIdx.x;
    const bool inb = p < n;
    const float4 pi = inb ? posid[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int id = inb ? pid_of(pi.w) : 0;
    const int oi = woff_of(pi.w);
    const bool act = inb && id >= R0.P.nb2;
    if (__ballot(act) == 0) return;   // wave-uniform
    with_rep<ENS>(R0, reps, rep_of, inb, id, [&](const GcRep& R) __attribute__((always_inline)) {
    const GcParams& P = R.P;
    const bool mov = act && id >= P.nf;
    const float ih = 1.0f / P.h;
    const float L0 = R.L[0], L1 = R.L[1], L2 = R.L[2];
    float d = 0.0f;  // self term added below
    float sw = 0.f, svx = 0.f, svy = 0.f, svz = 0.f;
    const int cnt = act ? nlc[p] : 0;
    int* wl = nl + (size_t)__builtin_amdgcn_readfirstlane(p >> 6) * (VL_ROWS * 64) + lane;
    auto take = [&](int j, float r2, int idj) {
        float q = __builtin_amdgcn_sqrtf(r2) * ih, om = 1.0f - q;
        float w = (1.0f + 3.0f * q) * om * om * om;
        if (mov) {
            d += idj >= P.nb2 ? w : 0.0f;
        } else if (idj >= P.nf) {
            float4 vj = velrho[j];
            sw += w; svx += w * vj.x; svy += w * vj.y; svz += w * vj.z;
        }
    };
    const int kmax = wave_max(cnt);
    // the codes of the next U entries are loaded while the current U
    // entries' positions are in flight (U = 8: 10M step 586 -> 504 us
    // against U = 4; U = 2 and U = 16 were slower)
    constexpr int U = 8;   // divides the 32 entries of a mask row
    int cn[U];
    auto load_codes = [&](int k) {
#pragma unroll
        for (int u = 0; u < U; ++u) cn[u] = cnt > 0 ? wl[64 * min(k + u, cnt - 1)] : 0;   // in bounds
    };
    if (kmax > 0) load_codes(0);
    for (int k0 = 0; k0 < kmax; k0 += 32) {
        unsigned m = 0u;
        const int k1 = min(k0 + 32, kmax);
        for (int k = k0; k < k1; k += U) {
            int c[U];
            float4 q[U];
#pragma unroll
            for (int u = 0; u < U; ++u) c[u] = cn[u];
#pragma unroll
            for (int u = 0; u < U; ++u) q[u] = posid[c[u] >> 6];
            if (k + U < kmax) load_codes(k + U);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool ok = k + u < cnt;
                const int oj = woff_of(q[u].w);
                float rx, ry, rz;
                if (__ballot(ok && ((c[u] & 63) | oi | oj))) {   // an image entry in the wave (x + 0 is x)
                    const float sx = eff_shift(c[u], oi, oj, 0, L0), sy = eff_shift(c[u], oi, oj, 1, L1),
                                sz = eff_shift(c[u], oi, oj, 2, L2);
                    rx = pdiff(pi.x, q[u].x, sx); ry = pdiff(pi.y, q[u].y, sy); rz = pdiff(pi.z, q[u].z, sz);
                } else {
                    rx = pi.x - q[u].x; ry = pi.y - q[u].y; rz = pi.z - q[u].z;
                }
                const float r2 = rx * rx + ry * ry + rz * rz;
                if (ok && r2 < P.h2) {
                    m |= 1u << (k + u - k0);
                    take(c[u] >> 6, r2, pid_of(q[u].w));
                }
            }
        }
        if (mov && cnt > k0) wl[64 * (VL_CAP + (k0 >> 5))] = (int)m;
    }
    if (act && cnt < 0) {   // the list overflowed at the rebuild
        int cx, cy, cz;
        key_to_cc(key[p], R, cx, cy, cz);
        visit2(R, offs, cx, cy, cz, [&](int jb, int je, float sx, float sy, float sz) {
            const int code = shift_code(sx, sy, sz);
            for (int j = jb; j < je; ++j) {
                const float4 pj = posid[j];
                const int oj = woff_of(pj.w);   // the cubes of the last rebuild, wraps since added
                const float ex = eff_shift(code, oi, oj, 0, L0), ey = eff_shift(code, oi, oj, 1, L1),
                            ez = eff_shift(code, oi, oj, 2, L2);
                const float rx = pdiff(pi.x, pj.x, ex), ry = pdiff(pi.y, pj.y, ey), rz = pdiff(pi.z, pj.z, ez);
                const float r2 = rx * rx + ry * ry + rz * rz;
                if (r2 < P.h2 && j != p) take(j, r2, pid_of(pj.w));
            }
        });
    }
    if (mov) {
        velrho[p].w = P.m * P.wnorm * (d + 1.0f);
    } else if (act && id < P.nf) {
        const float3 u = wall_velocity(P, pi.y, sw, svx, svy, svz);
        float4 v = velrho[p];
        v.x = u.x; v.y = u.y; v.z = u.z;
        velrho[p] = v;
    }
    });
}

// [18] wall velocity from the corrected velocities; walls' entry of vout
// (SK: skin mode, the cubes of the last rebuild with the wraps since.)
template <bool ENS, bool SK>
__global__ __launch_bounds__(BS) void k2_wallvel(int n, GcRep R0, const GcRep* __restrict__ reps,
    const int* __restrict__ rep_of, const float4* __restrict__ posid, float4* __restrict__ vout,
    const int* __restrict__ offs, const int* __restrict__ key)
{
    int p = xcd_block((n + BS - 1) / BS) * BS + threadIdx.x;
    if (p >= n) return;
    float4 pi = posid[p];
    int id = pid_of(pi.w);
    const int oi = SK ? woff_of(pi.w) : 0;
    if (id >= R0.P.nf) return;
    const GcRep& R = rep_of_id<ENS>(R0, reps, rep_of, id);
    const GcParams& P = R.P;
    if (id < P.nb2) { vout[p] = make_float4(0.f, 0.f, 0.f, P.rho0); return; }
    int cx, cy, cz;
    key_to_cc(key[p], R, cx, cy, cz);
    const float ih = 1.0f / P.h;
    float sw = 0.f, svx = 0.f, svy = 0.f, svz = 0.f;
    visit2(R, offs, cx, cy, cz, [&](int jb, int je, float sx, float sy, float sz) {
        const int code = shift_code(sx, sy, sz);
        for (int j = jb; j < je; ++j) {
            float4 pj = posid[j];
            if (SK) {
                const int oj = woff_of(pj.w);
                sx = eff_shift(code, oi, oj, 0, R.L[0]); sy = eff_shift(code, oi, oj, 1, R.L[1]);
                sz = eff_shift(code, oi, oj, 2, R.L[2]);
            }
            float rx = pdiff(pi.x, pj.x, sx), ry = pdiff(pi.y, pj.y, sy), rz = pdiff(pi.z, pj.z, sz);
            float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < P.h2 && pid_of(pj.w) >= P.nf) {
                float q = __builtin_amdgcn_sqrtf(r2) * ih, om = 1.0f - q;
                float w = (1.0f + 3.0f * q) * om * om * om;
                float4 vj = vout[j];
                sw += w; svx += w * vj.x; svy += w * vj.y; svz += w * vj.z;
            }
        }
    });
    const float3 u = wall_velocity(P, pi.y, sw, svx, svy, svz);
    vout[p] = make_float4(u.x, u.y, u.z, P.rho0);
}


// Unwrapped membrane vertex positions in id order (FP64, x[nm] y[nm] z[nm]):
// the membrane force kernel then reads a compact 24 B x Nm array instead of
// gathering coor / nthrobound / curpos from the whole slot-ordered state.
__global__ __launch_bounds__(BS) void k2_mgather(int nm, GcRep R0, const GcRep* __restrict__ reps,
    const int* __restrict__ cell_rep, const int* __restrict__ celltab,
    const int* __restrict__ curpos, const tfloat3* __restrict__ coor,
    const tint3* __restrict__ nthro, double* __restrict__ mpos)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nm) return;
    const GcRep& R = reps ? reps[cell_rep[celltab[v]]] : R0;
    int s = curpos[R0.P.nc + v];
    tfloat3 x = coor[s];
    tint3 nn = nthro[s];
    double L0 = (double)(R.bbox[1] - R.bbox[0]), L1 = (double)(R.bbox[3] - R.bbox[2]),
           L2 = (double)(R.bbox[5] - R.bbox[4]);
    mpos[v] = (double)x.x + (R.P.per[0] ? (double)nn.x * L0 : 0.0);
    mpos[nm + v] = (double)x.y + (R.P.per[1] ? (double)nn.y * L1 : 0.0);
    mpos[2 * nm + v] = (double)x.z + (R.P.per[2] ? (double)nn.z * L2 : 0.0);
}

}  // namespace

namespace {

struct Prof {
    hipEvent_t ev[32];
    const char* name[32];
    int n = 0;
};

void mark(Prof* pr, hipStream_t s, const char* name)
{
    if (!pr || pr->n >= 32) return;
    (void)hipEventRecord(pr->ev[pr->n], s);
    pr->name[pr->n] = name;
    pr->n++;
}

// Neighbour build: keys from the (optionally advanced) FP64 positions,
// counting sort by (cube, id), gather (+ predictor).
template <bool ENS>
int launch_build_t(gc_engine* e, int move, Prof* pr = nullptr)
{
    const int N = e->N, c = e->cur, nx = 1 - c;
    hipStream_t s = e->s;
    const int gb = xcd_grid(blocks_for(N, BS));
    const GcRep& R0 = e->hreps[0];
    // with membrane substepping the vertices are advanced by ms_step and
    // carry their end-of-step velocity (no predictor)
    const int mfree = e->ms == nullptr;
    {
        auto L = [&](auto kern, int tbs) {
            kern<<<xcd_grid(blocks_for(N, tbs)), tbs, 0, s>>>(N, R0, e->d_reps, e->d_rep_of, move, mfree,
                e->moving_walls,
                e->d_pos64[c], e->d_vout, e->d_force, e->d_sp[c], e->d_key0, e->d_rank, e->d_cnt);
        };
        switch (e->tune.v[TUNE_INTEGRATE]) {
            case 0: L(k2_integrate<ENS, 64>, 64); break;
            case 1: L(k2_integrate<ENS, 128>, 128); break;
            default: L(k2_integrate<ENS, 256>, 256); break;
        }
    }
    GC_LAUNCH_CHECK("engine integrate");
    mark(pr, s, "integrate+key");
    k2_tile_sums<<<e->ntile, CUBE_SCAN_BS, 0, s>>>(e->K, e->d_cnt, e->d_tsum, nullptr);
    k2_cube_offsets<<<e->ntile, CUBE_SCAN_BS, 0, s>>>(e->K, e->d_cnt, e->d_tsum, e->d_offs, e->d_ctr, move,
                                                      e->d_nlover, nullptr);
    GC_LAUNCH_CHECK("engine cube offsets");
    mark(pr, s, "cube offsets");
    k2_scatter<<<gb, BS, 0, s>>>(N, e->d_key0, e->d_rank, e->d_sp[c], e->d_offs, e->d_tk, e->d_tv,
                                 e->d_to, e->d_cnt, nullptr, nullptr, nullptr, nullptr);
    GC_LAUNCH_CHECK("engine scatter");
    k2_sortgather<ENS, false><<<gb, BS, 0, s>>>(N, R0, e->d_reps, e->d_rep_of, move, mfree, e->moving_walls,
        e->d_tk, e->d_tv,
        e->d_to, e->d_offs, e->d_curpos, e->d_pos64[c], e->d_pos64[nx], e->d_vout, e->d_force, e->d_nt[c],
        e->d_nt[nx], e->d_sp[nx], e->d_key, e->d_posid, e->d_coor, e->d_velrho, e->d_velhalf,
        e->moving_walls ? nullptr : e->d_wincl, e->d_act, e->d_fault, nullptr, nullptr, nullptr, nullptr,
        nullptr);
    GC_LAUNCH_CHECK("engine sort/gather");
    e->cur = nx;
    return GC_OK;
}

int debug_check_slot_maps(gc_engine* e);

int launch_build(gc_engine* e, int move)
{
    return e->ens() ? launch_build_t<true>(e, move) : launch_build_t<false>(e, move);
}

// [16] Morse (list iteration, or the cube scan for overflowed lists)
template <bool ENS>
void launch_agg(gc_engine* e)
{
    auto L = [&](auto kern) {
        kern<<<xcd_grid(blocks_for(e->nm, BS)), BS, 0, e->s>>>(e->nm, e->hreps[0], e->d_reps, e->d_rep_of,
            e->d_curpos, e->d_posid, e->d_offs, e->d_key, e->d_nl, e->d_nlc, e->d_celltab, e->d_force,
            e->d_aggforce, nl_stride(e));
    };
    if (e->skin_mode()) L(k2_agg<ENS, true>);
    else L(k2_agg<ENS, false>);
}

template <bool ENS>
int skin_lists_t(gc_engine* e, const int* gate, const int* rbflag, const unsigned char* gact);

// Skin mode: re-sort and list build of the replicas whose flag (rbflag)
// is set, all kernels returning at once unless *gate is set (decided on the
// device each step, or an all-ones flag for a host-driven rebuild).
template <bool ENS>
int skin_rebuild_t(gc_engine* e, const int* gate, const int* rbflag)
{
    const int N = e->N;
    hipStream_t s = e->s;
    const GcRep& R0 = e->hreps[0];
    const int mfree = e->ms == nullptr;
    const unsigned char* gact = ENS ? e->d_gact : nullptr;   // activity of the slot chunks
    if (ENS) {   // the chunks' activity; the counts of the other replicas persist
        const int nsc = chunks64(N), ncc = chunks64(e->K);
        k3_mark<<<gated_grid(blocks_for(nsc + ncc, BS), ENS), BS, 0, s>>>(nsc, ncc, gate, rbflag, e->d_chunkrep,
                                                                    e->d_gact, e->d_gact + act_bytes(N));
        k3_clear_counts<<<gated_grid(blocks_for(e->K, BS), ENS), BS, 0, s>>>(e->K, gate, rbflag, e->d_cubebase, e->nrep,
                                                                        e->d_gact + act_bytes(N), e->d_cnt);
    }
    k3_keys<ENS, 256><<<gated_grid(blocks_for(N, 256), ENS), 256, 0, s>>>(N, R0, e->d_reps, e->d_rep_of, gate, rbflag,
        gact, e->d_pos64[0], e->d_sp[0], e->d_key0, e->d_rank, e->d_cnt);
    k2_tile_sums<<<e->ntile, CUBE_SCAN_BS, 0, s>>>(e->K, e->d_cnt, e->d_tsum, gate);
    k2_cube_offsets<<<e->ntile, CUBE_SCAN_BS, 0, s>>>(e->K, e->d_cnt, e->d_tsum, e->d_offs, e->d_ctr, 0,
                                                      e->d_nlover, gate);
    // a batch re-sorts the rebuilding replicas only (the others keep their slots)
    const int* rbf = ENS ? rbflag : nullptr;
    k2_scatter<<<gated_grid(blocks_for(N, BS), ENS), BS, 0, s>>>(N, e->d_key0, e->d_rank, e->d_sp[0], e->d_offs,
                                                            e->d_tk, e->d_tv, e->d_to, e->d_cnt, gate,
                                                            e->d_rep_of, rbf, gact);
    k2_sortgather<ENS, true><<<gated_grid(blocks_for(N, BS), ENS), BS, 0, s>>>(N, R0, e->d_reps, e->d_rep_of, 0, mfree,
        e->moving_walls,
        e->d_tk, e->d_tv, e->d_to, e->d_offs, e->d_curpos, e->d_pos64[0], e->d_pos64[1], e->d_vout,
        e->d_force, e->d_nt[0], e->d_nt[1], e->d_sp[1], e->d_key, e->d_posid, e->d_coor, e->d_velrho,
        e->d_velhalf, e->moving_walls ? nullptr : e->d_wincl, e->d_act, e->d_fault, gate, rbflag, e->d_woff,
        e->d_sp[0], gact);
    k3_copyback<<<gated_grid(blocks_for(N, BS), ENS), BS, 0, s>>>(N, gate, e->d_pos64[1], e->d_pos64[0], e->d_nt[1],
                                                 e->d_nt[0], e->d_sp[1], e->d_sp[0], e->d_rep_of, rbf,
                                                 gact);
    GC_LAUNCH_CHECK("engine skin re-sort");
    return skin_lists_t<ENS>(e, gate, rbflag, gact);
}

// Skin mode: the Verlet lists of the replicas whose flag is set (gated).
template <bool ENS>
int skin_lists_t(gc_engine* e, const int* gate, const int* rbflag, const unsigned char* gact)
{
    const int N = e->N;
    hipStream_t s = e->s;
    const GcRep& R0 = e->hreps[0];
    {
        auto L = [&](auto kern, int tbs) {
            // a batch: a fixed grid (1024 groups' worth of 256 lanes) striding over the
            // groups (64 sheared replicas: list rebuild 283 us per step with 2048,
            // 253 with 1024, 308 with 512, 311 and 361 with 4096 and 8192)
            const int grid = ENS ? std::min(xcd_grid(blocks_for(N, tbs)), GATED_MAX_GROUPS / 2 * (256 / tbs))
                                 : xcd_grid(blocks_for(N, tbs));
            kern<<<grid, tbs, 0, s>>>(N, R0, e->d_reps, e->d_rep_of, e->d_posid,
                e->d_velrho, e->d_offs, e->d_key, e->d_nl, e->d_nlc, e->d_nlover, gate, rbflag, e->d_rbpos,
                gact);
        };
        switch (e->tune.v[TUNE_DENSITY]) {   // block size x LDS row stage
            case 0: L(k2_density<ENS, 64, 0, true>, 64); break;
            case 1: L(k2_density<ENS, 64, 128, true>, 64); break;
            case 2: L(k2_density<ENS, 64, 192, true>, 64); break;
            case 3: L(k2_density<ENS, 64, 320, true>, 64); break;
            case 4: L(k2_density<ENS, 128, 0, true>, 128); break;
            case 5: L(k2_density<ENS, 128, 128, true>, 128); break;
            case 6: L(k2_density<ENS, 128, 192, true>, 128); break;
            case 7: L(k2_density<ENS, 128, 320, true>, 128); break;
            case 8: L(k2_density<ENS, 256, 0, true>, 256); break;
            case 9: L(k2_density<ENS, 256, 128, true>, 256); break;
            case 11: L(k2_density<ENS, 256, 320, true>, 256); break;
            default: L(k2_density<ENS, 256, 192, true>, 256); break;
        }
    }
    GC_LAUNCH_CHECK("engine list build");
    return GC_OK;
}

// Skin mode, after a host-driven build into buffer 1 (creation, SetState):
// buffer 0 made current, every replica's lists built, the displacement
// bounds cleared.
int skin_init(gc_engine* e)
{
    if (!e->skin_mode()) return GC_OK;
    const int N = e->N;
    hipStream_t s = e->s;
    if (e->cur == 1) {
        k3_copyback<<<gated_grid(blocks_for(N, BS)), BS, 0, s>>>(N, e->d_ones, e->d_pos64[1], e->d_pos64[0], e->d_nt[1],
                                                     e->d_nt[0], e->d_sp[1], e->d_sp[0], nullptr, nullptr, nullptr);
        GC_LAUNCH_CHECK("engine skin copy");
        e->cur = 0;
    }
    if (e->ens()) {   // the cube counts persist between a batch's rebuilds
        k3_counts<<<blocks_for(e->K, BS), BS, 0, s>>>(e->K, e->d_offs, e->d_cnt);
        GC_LAUNCH_CHECK("engine skin counts");
    }
    int st = e->ens() ? skin_lists_t<true>(e, e->d_ones, e->d_ones, nullptr)
                      : skin_lists_t<false>(e, e->d_ones, e->d_ones, nullptr);
    if (st) return st;
    const int nr = e->nrep;
    if ((st = hip_check(hipMemsetAsync(e->d_woff, 0, (size_t)N, s), "skin init")) ||
        (st = hip_check(hipMemsetAsync(e->d_disp2, 0, sizeof(unsigned) * DSTAT_N * nr, s), "skin init")) ||
        (st = hip_check(hipMemsetAsync(e->d_bmax, 0, sizeof(float) * skin_bmax_size(e), s), "skin init")) ||
        (st = hip_check(hipMemsetAsync(e->d_rbflag, 0, sizeof(int) * nr, s), "skin init")))
        return st;
    return hip_check(hipStreamSynchronize(s), "skin init sync");
}

// Skin mode, host-driven builds (creation, SetState): the given positions
// may lie outside a periodic face (the per-step scheme wraps them in its
// first step).  The build then runs on the wrapped positions (the lists'
// frame, in `frame`, id order, SoA), and afterwards the stored positions are
// put back unchanged with their offset to that frame as the wrap code (the
// first step's advance wraps them and clears it).  pos: id order, SoA.
void skin_frame(const gc_engine* e, const std::vector<int>& rep_of, const std::vector<double>& pos,
                std::vector<double>& frame, std::vector<unsigned char>& code)
{
    const size_t N = (size_t)e->N;
    frame = pos;
    code.assign(N, 0);
    for (size_t i = 0; i < N; ++i) {
        const GcRep& R = e->hreps[rep_of.empty() ? 0 : rep_of[i]];
        int k[3] = {0, 0, 0};
        for (int d = 0; d < 3; ++d) {
            if (!R.P.per[d]) continue;
            double& u = frame[(size_t)d * N + i];
            const double lo = R.bbox[2 * d], hi = R.bbox[2 * d + 1];
            const double L = (double)(R.bbox[2 * d + 1] - R.bbox[2 * d]);
            if (u < lo) { u += L; k[d] = 1; }          // frame = stored + L
            else if (u >= hi) { u -= L; k[d] = -1; }   // frame = stored - L
        }
        auto enc = [](int v) { return v < 0 ? 1 : (v > 0 ? 2 : 0); };   // image code: stored + k L = frame
        code[i] = (unsigned char)(enc(k[0]) | (enc(k[1]) << 2) | (enc(k[2]) << 4));
    }
}

// After the build and the lists: the stored positions back in slot order,
// their FP32 copies and wrap codes.
int skin_restore_stored(gc_engine* e, const std::vector<double>& pos, const std::vector<unsigned char>& code)
{
    const int N = e->N;
    bool any = false;
    for (unsigned char c : code) any = any || c;
    if (!any) return GC_OK;   // every position was inside the box: nothing to put back
    std::vector<int> sp(N);
    int st = hip_check(hipStreamSynchronize(e->s), "skin frame sync");
    if (!st) st = hip_check(hipMemcpy(sp.data(), e->d_sp[0], sizeof(int) * N, hipMemcpyDeviceToHost), "skin frame D2H");
    if (st) return st;
    std::vector<double> p64(3 * (size_t)N);
    std::vector<float4> pid(N);
    std::vector<tfloat3> co(N);
    std::vector<unsigned char> wc(N);
    for (int s = 0; s < N; ++s) {
        const int id = sp[s];
        const double x = pos[id], y = pos[(size_t)N + id], z = pos[2 * (size_t)N + id];
        p64[s] = x; p64[(size_t)N + s] = y; p64[2 * (size_t)N + s] = z;
        const int w = id | ((int)code[id] << 25);
        float wf;
        memcpy(&wf, &w, sizeof(wf));
        pid[s] = make_float4((float)x, (float)y, (float)z, wf);
        co[s] = tfloat3{(float)x, (float)y, (float)z};
        wc[s] = code[id];
    }
    if ((st = upload(e->d_pos64[0], p64.data(), 3 * (size_t)N, e->s)) ||
        (st = upload(e->d_posid, pid.data(), (size_t)N, e->s)) ||
        (st = upload(e->d_coor, co.data(), (size_t)N, e->s)) ||
        (st = upload(e->d_woff, wc.data(), (size_t)N, e->s)))
        return st;
    return hip_check(hipStreamSynchronize(e->s), "skin frame sync");
}

// Skin mode: [1,2,5] advance and predictor, the rebuild decision, the
// rebuild (gated), [7,9] density over the lists, [10-13,17] force over the
// neighbour bits.
template <bool ENS>
int skin_fluid_t(gc_engine* e, int driveon, Prof* pr)
{
    const GcRep& R0 = e->hreps[0];
    const int N = e->N;
    hipStream_t s = e->s;
    const int mfree = e->ms == nullptr;
    int nb_adv = 0, adv_bs = 0;
    {
        auto L = [&](auto kern, int tbs) {
            kern<<<xcd_grid(blocks_for(N, tbs)), tbs, 0, s>>>(N, R0, e->d_reps, e->d_rep_of, mfree,
                e->moving_walls, e->d_pos64[0], e->d_vout, e->d_force, e->d_sp[0], e->d_nt[0], e->d_woff,
                e->d_rbpos, e->d_posid, e->d_coor, e->d_velrho, e->d_velhalf, e->d_bmax, e->d_brep, e->d_disp2);
            nb_adv = blocks_for(N, tbs);
            adv_bs = tbs;
        };
        switch (e->tune.v[TUNE_INTEGRATE]) {
            case 0: L(k3_advance<ENS, 64>, 64); break;
            case 1: L(k3_advance<ENS, 128>, 128); break;
            default: L(k3_advance<ENS, 256>, 256); break;
        }
    }
    // the substeps of the previous step wrote their vertices' maxima after
    // the advance's region (membrane substepping: one suspension)
    const int nb_ms = e->ms ? blocks_for(e->nm, BS) : 0;
    if (e->ens() && nb_ms == 0)
        k3_reduce_rep<<<e->nrep, REDUCE_BS, 0, s>>>(e->d_bmax, e->d_brep, e->d_slotbase, adv_bs, e->d_disp2);
    else
        k3_reduce<<<REDUCE_NB, REDUCE_BS, 0, s>>>(e->nrep, e->d_bmax, e->d_brep, nb_adv, skin_ms_off(e), nb_ms,
                                                  e->d_disp2);
    k3_decide<<<1, 64, 0, s>>>(e->nrep, e->d_reps, e->d_disp2, e->d_rbflag, e->d_rbany, e->d_nrebuild, e->d_ctr);
    GC_LAUNCH_CHECK("engine advance");
    mark(pr, s, "advance+predict");
    int st = skin_rebuild_t<ENS>(e, e->d_rbany, e->d_rbflag);
    if (st) return st;
    if ((st = debug_check_slot_maps(e))) return st;
    mark(pr, s, "list rebuild");
    {
        auto L = [&](auto kern, int tbs) {
            kern<<<xcd_grid(blocks_for(N, tbs)), tbs, 0, s>>>(N, R0, e->d_reps, e->d_rep_of, e->d_posid,
                e->d_velrho, e->d_offs, e->d_key, e->d_nl, e->d_nlc);
        };
        const int v = e->tune.v[TUNE_DENSITY];
        if (v < 4) L(k3_density<ENS, 64>, 64);
        else if (v < 8) L(k3_density<ENS, 128>, 128);
        else L(k3_density<ENS, 256>, 256);
    }
    GC_LAUNCH_CHECK("engine density (lists)");
    mark(pr, s, "density+wall velocity");
    {
        auto L = [&](auto kern, int tbs) {
            if (e->nmov > 0)
                kern<<<xcd_grid(blocks_for(e->nmov, tbs)), tbs, 0, s>>>(e->nmov, R0, e->d_reps, e->d_rep_of,
                    driveon, e->d_act, e->d_posid, e->d_velrho, e->d_offs, e->d_key, e->d_nl, e->d_nlc,
                    e->d_velhalf, e->d_force, e->d_vout, e->d_ctr, e->d_fault);
        };
        switch (e->tune.v[TUNE_FORCE]) {
            case 0: L(k2_force<ENS, 64, true>, 64); break;
            case 1: L(k2_force<ENS, 128, true>, 128); break;
            default: L(k2_force<ENS, 256, true>, 256); break;
        }
    }
    GC_LAUNCH_CHECK("engine force");
    return GC_OK;
}

template <bool ENS>
int launch_step_t(gc_engine* e, int driveon, Prof* pr)
{
    const GcRep& R0 = e->hreps[0];
    const int N = e->N;
    hipStream_t s = e->s;
    mark(pr, s, "start");
    int st;
    if (e->skin_mode()) {
        if ((st = skin_fluid_t<ENS>(e, driveon, pr))) return st;
    } else {
    if ((st = launch_build_t<ENS>(e, 1, pr))) return st;
    if ((st = debug_check_slot_maps(e))) return st;
    mark(pr, s, "sort+gather+predict");
    {
        auto L = [&](auto kern, int tbs) {
            kern<<<xcd_grid(blocks_for(N, tbs)), tbs, 0, s>>>(N, R0, e->d_reps, e->d_rep_of, e->d_posid,
                e->d_velrho, e->d_offs, e->d_key, e->d_nl, e->d_nlc, e->d_nlover, nullptr, nullptr, nullptr,
                nullptr);
        };
        switch (e->tune.v[TUNE_DENSITY]) {   // block size x LDS row stage
            case 0: L(k2_density<ENS, 64, 0>, 64); break;
            case 1: L(k2_density<ENS, 64, 128>, 64); break;
            case 2: L(k2_density<ENS, 64, 192>, 64); break;
            case 3: L(k2_density<ENS, 64, 320>, 64); break;
            case 4: L(k2_density<ENS, 128, 0>, 128); break;
            case 5: L(k2_density<ENS, 128, 128>, 128); break;
            case 6: L(k2_density<ENS, 128, 192>, 128); break;
            case 7: L(k2_density<ENS, 128, 320>, 128); break;
            case 8: L(k2_density<ENS, 256, 0>, 256); break;
            case 9: L(k2_density<ENS, 256, 128>, 256); break;
            case 11: L(k2_density<ENS, 256, 320>, 256); break;
            default: L(k2_density<ENS, 256, 192>, 256); break;
        }
    }
    GC_LAUNCH_CHECK("engine density");
    mark(pr, s, "density+wall velocity");
    {
        auto L = [&](auto kern, int tbs) {
            if (e->nmov > 0)
                kern<<<xcd_grid(blocks_for(e->nmov, tbs)), tbs, 0, s>>>(e->nmov, R0, e->d_reps, e->d_rep_of,
                    driveon, e->d_act, e->d_posid, e->d_velrho, e->d_offs, e->d_key, e->d_nl, e->d_nlc,
                    e->d_velhalf, e->d_force, e->d_vout, e->d_ctr, e->d_fault);
        };
        switch (e->tune.v[TUNE_FORCE]) {
            case 0: L(k2_force<ENS, 64>, 64); break;
            case 1: L(k2_force<ENS, 128>, 128); break;
            default: L(k2_force<ENS, 256>, 256); break;
        }
    }
    GC_LAUNCH_CHECK("engine force");
    }
    mark(pr, s, "sdpd force");
    const int c = e->cur;
    if (e->ncell > 0 && e->ms) {
        if (e->any_morse) {
            launch_agg<ENS>(e);
            GC_LAUNCH_CHECK("engine morse");
        }
        mark(pr, s, "morse");
        if ((st = ms_step(e))) return st;
        mark(pr, s, "membrane substeps");
    } else if (e->ncell > 0) {
        const GcParams& P = R0.P;
        if (e->any_morse) {
            launch_agg<ENS>(e);
            GC_LAUNCH_CHECK("engine morse");
        }
        mark(pr, s, "morse");
        DefArgs a;
        a.edge = e->d_edge; a.tri = e->d_tri; a.coor = e->d_coor; a.nthro = e->d_nt[c];
        a.curpos = e->d_curpos; a.lmax = e->d_lmax; a.mexp = e->d_mexp; a.p = e->d_p;
        a.kp = e->d_kp; a.th0 = e->d_th0; a.cth0 = e->d_cth0; a.sth0 = e->d_sth0;
        a.kb = e->d_kb; a.reftriarea = e->d_reftriarea;
        a.kd = e->d_kd; a.ka = e->d_ka; a.Ac = e->d_area_para; a.A0c = e->d_refarea_para;
        a.kv = e->d_kv; a.Vc = e->d_vol_para; a.V0c = e->d_refvol_para;
        a.es = e->d_es; a.ee = e->d_ee; a.ts = e->d_ts; a.te = e->d_te;
        a.defforce = e->d_defforce; a.force = e->d_force;
        a.velhalf = e->d_velhalf; a.velrho = e->d_velrho; a.vout = e->d_vout;
        a.nm = e->nm;
        a.fault = e->d_fault;
        if (ENS) { a.reps = e->d_reps; a.cell_rep = e->d_cell_rep; a.celltab = e->d_celltab; }
        if (use_cell_kernel(e)) {
            // geometry, membrane force and corrector in one workgroup per cell
            mark(pr, s, "membrane geometry");
            k_cell_membrane<<<e->ncell, GC_CELL_BS, 0, s>>>(P, e->d_bbox, cell_args(e, e->step, 0), a);
            GC_LAUNCH_CHECK("engine membrane (per cell)");
        } else {
            k_geo_cell<<<e->ncell, GC_GEO_BS, 0, s>>>(P, e->step, ENS ? e->d_reps : nullptr,
                ENS ? e->d_cell_rep : nullptr, e->d_bbox, e->d_curpos, e->d_tri,
                e->d_betri, e->d_nt[c], e->d_coor, e->d_trinor, e->d_triarea, e->d_reftriarea,
                e->d_tricen, e->d_refarea_para, e->d_refvol_para, e->d_area_para, e->d_vol_para,
                e->d_area, e->d_vol);
            GC_LAUNCH_CHECK("engine geo");
            mark(pr, s, "membrane geometry");
            k2_mgather<<<blocks_for(e->nm, BS), BS, 0, s>>>(e->nm, R0, ENS ? e->d_reps : nullptr,
                e->d_cell_rep, e->d_celltab, e->d_curpos, e->d_coor, e->d_nt[c], e->d_mpos);
            GC_LAUNCH_CHECK("engine membrane gather");
            a.mpos = e->d_mpos;
            if ((st = membrane_force(e, P, a, s))) return st;
        }
        mark(pr, s, "membrane force+corrector");
    }
    // [18] CalBound2Vel after the corrector only feeds the output: the next
    // step overwrites the walls' velocity (k2_density) before anything reads
    // it.  It is computed when the state leaves the engine (finish_walls).
    e->walls_dirty = e->any_walls;
    e->step += 1;
    return GC_OK;
}

}  // namespace

CellArgs gc::cell_args(gc_engine* e, long step, int round_geo)
{
    CellArgs ca;
    ca.range = e->d_cellr; ca.betri = e->d_betri; ca.step = step; ca.round_geo = round_geo;
    ca.trinor = e->d_trinor; ca.triarea = e->d_triarea; ca.reftriarea = e->d_reftriarea;
    ca.tricen = e->d_tricen; ca.refarea_para = e->d_refarea_para; ca.refvol_para = e->d_refvol_para;
    ca.area_para = e->d_area_para; ca.vol_para = e->d_vol_para; ca.area = e->d_area; ca.vol = e->d_vol;
    return ca;
}

int gc::membrane_force(gc_engine* e, const GcParams& P, const DefArgs& a, hipStream_t s)
{
    // 64-thread groups: a few hundred RBCs give only ~Nm/256 groups of 256
    const int v = e->tune.v[TUNE_MEMBRANE];
    const int dg = 64 << (v < 3 ? v : 0);   // 64, 128 or 256 threads
    k_edge_force<<<xcd_grid(blocks_for(e->nedge, dg)), dg, 0, s>>>(P, e->nedge, e->d_bbox, a, e->d_ef);
    k_tri_force<<<xcd_grid(blocks_for(e->ntri, dg)), dg, 0, s>>>(P, e->ntri, e->d_bbox, a, e->d_tf);
    k_vert_force<<<xcd_grid(blocks_for(e->nm, dg)), dg, 0, s>>>(P, e->nm, a, e->d_ef, e->d_tf);
    GC_LAUNCH_CHECK("engine membrane force");
    return GC_OK;
}

namespace {

int launch_step(gc_engine* e, int driveon, Prof* pr)
{
    return e->ens() ? launch_step_t<true>(e, driveon, pr) : launch_step_t<false>(e, driveon, pr);
}

// Wall velocities from the corrected flow (UpdatePar's final CalBound2Vel,
// Particle.cu:1327) for the exported state.
int finish_walls(gc_engine* e)
{
    if (!e->walls_dirty) return GC_OK;
    const int gb = xcd_grid(blocks_for(e->N, BS));
    auto L = [&](auto kern) {
        kern<<<gb, BS, 0, e->s>>>(e->N, e->hreps[0], e->d_reps, e->d_rep_of, e->d_posid, e->d_vout, e->d_offs,
                                  e->d_key);
    };
    if (e->ens()) { if (e->skin_mode()) L(k2_wallvel<true, true>); else L(k2_wallvel<true, false>); }
    else if (e->skin_mode()) L(k2_wallvel<false, true>);
    else L(k2_wallvel<false, false>);
    GC_LAUNCH_CHECK("engine wall velocity");
    e->walls_dirty = false;
    return GC_OK;
}

void destroy(gc_engine* e)
{
    if (!e) return;
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->s) (void)hipStreamSynchronize(e->s);
    if (e->ms) ms_destroy(e->ms);
    for (void* p : e->allocs) (void)hipFree(p);
    if (e->s) (void)hipStreamDestroy(e->s);
    delete e;
}

// After a build without predictor (create / SetState) the arrays the gather
// does not carry (vout, force) are still in the previous slot order: put
// them, and velhalf, into the new order on the host.
int reorder_carried(gc_engine* e, const std::vector<float4>& vout_id,
                    const std::vector<tfloat3>& force_id, const std::vector<tfloat3>& vh_id)
{
    const int N = e->N;
    std::vector<int> sp(N);
    int st = hip_check(hipStreamSynchronize(e->s), "reorder sync");
    if (st) return st;
    if ((st = hip_check(hipMemcpy(sp.data(), e->d_sp[e->cur], sizeof(int) * N, hipMemcpyDeviceToHost), "reorder sp")))
        return st;
    std::vector<float4> vo(N);
    std::vector<tfloat3> f(N), vh(N);
    for (int s = 0; s < N; ++s) { vo[s] = vout_id[sp[s]]; f[s] = force_id[sp[s]]; vh[s] = vh_id[sp[s]]; }
    if ((st = upload(e->d_vout, vo.data(), N, e->s)) || (st = upload(e->d_force, f.data(), N, e->s)) ||
        (st = upload(e->d_velhalf, vh.data(), N, e->s)))
        return st;
    return hip_check(hipStreamSynchronize(e->s), "reorder sync2");
}

// walls per cube from the current keys (walls never move), the inclusive
// prefix over cubes, and the number of moving particles
int wall_tables(gc_engine* e)
{
    const int N = e->N, K = e->K, nf = e->hreps[0].P.nf;
    std::vector<int> sp(N), key(N), w(K + 1, 0);
    int st = hip_check(hipStreamSynchronize(e->s), "wall table sync");
    if (!st) st = hip_check(hipMemcpy(sp.data(), e->d_sp[e->cur], sizeof(int) * N, hipMemcpyDeviceToHost), "wall table");
    if (!st) st = hip_check(hipMemcpy(key.data(), e->d_key, sizeof(int) * N, hipMemcpyDeviceToHost), "wall table");
    if (st) return st;
    int nwall = 0;
    for (int s = 0; s < N; ++s)
        if (sp[s] < nf) {
            if (key[s] < 0 || key[s] >= K) return set_error(GC_ERR_GRID, "wall table: key out of range");
            w[key[s]] += 1;
            ++nwall;
        }
    for (int k = 1; k < K; ++k) w[k] += w[k - 1];
    // moving walls change cubes: the force pass then runs over every slot
    e->nmov = e->moving_walls ? N : N - nwall;
    if ((st = upload(e->d_wincl, w.data(), (size_t)K, e->s))) return st;
    std::vector<int> act;
    if (e->skin_mode()) {
        // the build that preceded this table wrote the moving-slot map
        // without it, and skin mode re-sorts only at a list rebuild: write
        // the map here (the moving particles' slots in slot order)
        act.reserve((size_t)e->nmov);
        for (int s = 0; s < N; ++s)
            if (e->moving_walls || sp[s] >= nf) act.push_back(s);
        if ((st = upload(e->d_act, act.data(), act.size(), e->s))) return st;
    }
    return hip_check(hipStreamSynchronize(e->s), "wall table sync");   // w, act are freed on return
}

int validate_desc(const gc_case_desc* d)
{
    if (!d || !d->coor || d->nBound1Par < 0 || d->nBound2Par < 0 || d->nFluidPar < 0 ||
        d->nCell < 0 || d->nCellPar < 0)
        return set_error(GC_ERR_INVALID, "EngineCreate: bad descriptor");
    if (d->nCell > 0 && (!d->tri || !d->edge || !d->celltab || !d->beginendtri ||
                         !d->shearmod_para || !d->bendingmod_para || !d->hlarea_para ||
                         !d->hgarea_para || !d->hvol_para || d->nAllTri <= 0 ||
                         d->nAllEdge <= 0 || d->nCellPar <= 0))
        return set_error(GC_ERR_INVALID, "EngineCreate: membrane arrays missing");
    const Structconst& c = d->consts;
    if (!(c.CutRadius > 0) || !(c.Mass > 0) || !(c.TimeStep > 0) || !(c.DenInit > 0))
        return set_error(GC_ERR_INVALID, "EngineCreate: CutRadius, Mass, TimeStep, DenInit must be > 0");
    if (5.0 * c.ShearVis / 3.0 - c.BulkVis < 0)
        return set_error(GC_ERR_INVALID, "EngineCreate: need 5*ShearVis/3 >= BulkVis");
    if (c.surEnergy != 0 && c.MaxCut > c.CutRadius)
        return set_error(GC_ERR_GRID, "EngineCreate: MaxCut must not exceed CutRadius");
    if (!(d->ext.wlc_ratio > 1.0f))
        return set_error(GC_ERR_INVALID, "EngineCreate: wlc_ratio must be > 1");
    if (!(d->skin == 0.0f || (d->skin >= 0.01f && d->skin <= 10.0f * c.CutRadius)))
        return set_error(GC_ERR_INVALID, "EngineCreate: skin must be 0 or in [0.01, 10 CutRadius]");
    for (int cc = 0; cc < d->nCell; ++cc)
        if (d->beginendtri[cc].x < 0 || d->beginendtri[cc].y >= d->nAllTri ||
            d->beginendtri[cc].x > d->beginendtri[cc].y)
            return set_error(GC_ERR_TOPOLOGY, "EngineCreate: bad BeginEndTri");
    for (int v = 0; v < d->nCellPar && d->nCell > 0; ++v)
        if (d->celltab[v] < 0 || d->celltab[v] >= d->nCell)
            return set_error(GC_ERR_TOPOLOGY, "EngineCreate: CellTab out of range");
    return GC_OK;
}

gc_engine* create(const gc_case_desc* descs, int nrep, int* status)
{
    int dummy;
    if (!status) status = &dummy;
    *status = GC_OK;
    if (!descs || nrep < 1) { *status = set_error(GC_ERR_INVALID, "EngineCreate: no case"); return nullptr; }
    for (int r = 0; r < nrep; ++r)
        if ((*status = validate_desc(&descs[r]))) return nullptr;
    for (int r = 1; r < nrep; ++r)
        if ((descs[r].skin > 0.0f) != (descs[0].skin > 0.0f)) {
            *status = set_error(GC_ERR_INVALID, "EnsembleCreate: replicas must all use a skin or none");
            return nullptr;
        }
    for (int r = 1; r < nrep; ++r)
        if (descs[r].ext.noise_counter != descs[0].ext.noise_counter) {
            *status = set_error(GC_ERR_INVALID, "EnsembleCreate: replicas must share the noise counter");
            return nullptr;
        }
    gc_engine* e = new gc_engine();
    int st = GC_OK;
    auto fail = [&](int code) { *status = code; destroy(e); return (gc_engine*)nullptr; };
    e->nrep = nrep;
    e->meta.resize(nrep);
    e->hreps.resize(nrep);
    // global id layout and totals
    long tb1 = 0, tb2 = 0, tfl = 0, tm = 0, tc = 0, tt = 0, te = 0;
    for (int r = 0; r < nrep; ++r) {
        const gc_case_desc& d = descs[r];
        RepMeta& m = e->meta[r];
        m.nb1 = d.nBound1Par; m.nb2 = d.nBound2Par; m.nfl = d.nFluidPar;
        m.ncell = d.nCell; m.nm = d.nCell > 0 ? d.nCellPar : 0;
        m.ntri = d.nCell > 0 ? d.nAllTri : 0; m.nedge = d.nCell > 0 ? d.nAllEdge : 0;
        m.vtx_off = (int)tm; m.cell_off = (int)tc; m.tri_off = (int)tt; m.edge_off = (int)te;
        tb1 += m.nb1; tb2 += m.nb2; tfl += m.nfl; tm += m.nm; tc += m.ncell; tt += m.ntri; te += m.nedge;
        Structconst c = d.consts;
        c.nBound2Start = (unsigned)m.nb1;
        c.nFluidStart = (unsigned)(m.nb1 + m.nb2);
        c.nCellStart = c.nFluidStart + (unsigned)m.nfl;
        c.nCellPar = (unsigned)m.nm;
        c.nTotalPar = c.nCellStart + c.nCellPar;
        m.consts = c;
    }
    const long Ntot = tb1 + tb2 + tfl + tm;
    if (Ntot >= (1L << 25))  // the force kernel packs slot << 6 into an int
        return fail(set_error(GC_ERR_INVALID, "EngineCreate: more than 2^25 particles"));
    e->N = (int)Ntot; e->nm = (int)tm; e->ncell = (int)tc; e->ntri = (int)tt; e->nedge = (int)te;
    e->consts = e->meta[0].consts;
    e->consts.nBound2Start = (unsigned)tb1;
    e->consts.nFluidStart = (unsigned)(tb1 + tb2);
    e->consts.nCellStart = (unsigned)(tb1 + tb2 + tfl);
    e->consts.nCellPar = (unsigned)tm;
    e->consts.nTotalPar = (unsigned)Ntot;
    e->ext = descs[0].ext;
    // per-replica tables
    long koff = 0;
    long b1o = 0, b2o = tb1, flo = tb1 + tb2, mo = tb1 + tb2 + tfl;
    for (int r = 0; r < nrep; ++r) {
        const gc_case_desc& d = descs[r];
        RepMeta& m = e->meta[r];
        GcRep& R = e->hreps[r];
        memset(&R, 0, sizeof(R));
        Structconst cg = d.consts;  // the replica's physics with the global thresholds
        cg.nBound2Start = e->consts.nBound2Start; cg.nFluidStart = e->consts.nFluidStart;
        cg.nCellStart = e->consts.nCellStart; cg.nCellPar = e->consts.nCellPar;
        cg.nTotalPar = e->consts.nTotalPar;
        make_params(cg, d.ext, &R.P);
        if (d.ext.wall_u_lo != 0.0f || d.ext.wall_u_hi != 0.0f) e->moving_walls = 1;
        if (r == 0) {
            const char* cs = getenv("GPUCELL_CHECK_SLOTS");
            e->check_slots = cs && cs[0] == '1';
        }
        int Kr = 0;
        float outer[3];
        // skin mode: cubes wider than the list radius h + skin (with a margin
        // for the FP32 key rounding), padded by their width on walled axes
        const float cut = d.skin > 0.0f ? (d.consts.CutRadius + d.skin) * 1.0001f + 1.0e-3f : d.consts.CutRadius;
        if ((st = gc_ConfigDomain(d.bbox, cut, d.ext.periodic, R.nc, outer, R.invw, &Kr)))
            return fail(st);
        if (d.skin > 0.0f) {
            for (int q = 0; q < 3; ++q) R.P.off[q] = d.ext.periodic[q] ? 0.0f : cut;
            const float rl = d.consts.CutRadius + d.skin;
            R.list_r2 = rl * rl;
            // FP32 positions and distances carry rounding errors well below 1e-3
            R.half_skin = 0.5f * d.skin - 1.0e-3f;
        }
        for (int q = 0; q < 3; ++q) R.L[q] = outer[q];
        memcpy(R.bbox, d.bbox, sizeof(R.bbox));
        R.cube_off = (int)koff;
        R.tys = cube_tile_shift(R.nc[0], R.nc[1]);
        koff += Kr;
        R.id_off[0] = (int)b1o; R.id_off[1] = (int)b2o; R.id_off[2] = (int)flo; R.id_off[3] = (int)mo;
        R.loc_start[0] = 0; R.loc_start[1] = m.nb1; R.loc_start[2] = m.nb1 + m.nb2;
        R.loc_start[3] = m.nb1 + m.nb2 + m.nfl;
        b1o += m.nb1; b2o += m.nb2; flo += m.nfl; mo += m.nm;
        e->any_morse = e->any_morse || (d.consts.surEnergy != 0.0f && m.ncell > 0);
        e->any_walls = e->any_walls || (m.nb1 + m.nb2 > 0);
        e->any_bound1 = e->any_bound1 || m.nb1 > 0;
    }
    if (koff > (1L << 30)) return fail(set_error(GC_ERR_GRID, "EngineCreate: too many cubes"));
    e->K = (int)koff;
    e->skin = descs[0].skin > 0.0f ? descs[0].skin : 0.0f;
    e->nlcap = e->skin_mode() ? VL_CAP : NL_CAP;
    e->ntile = (e->K + CUBE_TILE - 1) / CUBE_TILE;
    if ((st = hip_check(hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking), "engine stream")))
        return fail(st);

    const size_t N = (size_t)e->N, K = (size_t)e->K;
    if ((st = e->alloc(&e->d_reps, (size_t)nrep)) || (st = e->alloc(&e->d_rep_of, N)) ||
        (st = e->alloc(&e->d_cell_rep, (size_t)(e->ncell > 0 ? e->ncell : 1))) ||
        (st = e->alloc(&e->d_bbox, 6)))
        return fail(st);
    for (int b = 0; b < 2; ++b)
        if ((st = e->alloc(&e->d_pos64[b], 3 * N)) || (st = e->alloc(&e->d_nt[b], N)) ||
            (st = e->alloc(&e->d_sp[b], N)))
            return fail(st);
    if ((st = e->alloc(&e->d_posid, N)) || (st = e->alloc(&e->d_velrho, N)) ||
        (st = e->alloc(&e->d_vout, N)) || (st = e->alloc(&e->d_coor, N)) ||
        (st = e->alloc(&e->d_velhalf, N)) || (st = e->alloc(&e->d_force, N)) ||
        (st = e->alloc(&e->d_key0, N)) || (st = e->alloc(&e->d_key, N)) ||
        (st = e->alloc(&e->d_rank, N)) || (st = e->alloc(&e->d_cnt, K + 1)) ||
        (st = e->alloc(&e->d_offs, K + 1)) || (st = e->alloc(&e->d_tk, N)) ||
        (st = e->alloc(&e->d_tv, N)) || (st = e->alloc(&e->d_to, N)) || (st = e->alloc(&e->d_curpos, N)) ||
        (st = e->alloc(&e->d_nl, (size_t)nl_stride(e) * (((size_t)N + 63) / 64))) || (st = e->alloc(&e->d_nlc, N)) ||
        (st = e->alloc(&e->d_ctr, 1)) || (st = e->alloc(&e->d_fault, 1)) || (st = e->alloc(&e->d_nlover, 1)) ||
        (st = hip_check(hipMemset(e->d_fault, 0, sizeof(int)), "engine fault word")) ||
        (st = hip_check(hipMemset(e->d_nlover, 0, sizeof(int)), "engine overflow counter")) ||
        (st = e->alloc(&e->d_tsum, (size_t)e->ntile)) || (st = e->alloc(&e->d_act, N)) ||
        (st = e->alloc(&e->d_wincl, K + 1)) ||
        (st = hip_check(hipMemset(e->d_wincl, 0, sizeof(int) * (K + 1)), "engine wall table")) ||
        (st = hip_check(hipMemset(e->d_act, 0, sizeof(int) * N), "engine moving slots")))
        return fail(st);
    if (e->skin_mode()) {
        const size_t nr = (size_t)nrep;
        std::vector<int> ones(nr, 1), zero(nr, 0);
        if ((st = e->alloc(&e->d_disp2, DSTAT_N * nr)) || (st = e->alloc(&e->d_rbpos, (size_t)N)) ||
            (st = e->alloc(&e->d_bmax, skin_bmax_size(e))) || (st = e->alloc(&e->d_brep, skin_bmax_size(e))) ||
            (st = e->alloc(&e->d_rbflag, nr)) || (st = e->alloc(&e->d_rbany, 1)) ||
            (st = e->alloc(&e->d_nrebuild, nr)) || (st = e->alloc(&e->d_ones, nr)) ||
            (st = e->alloc(&e->d_woff, (size_t)N)) ||
            (st = hip_check(hipMemset(e->d_woff, 0, (size_t)N), "engine wrap codes")) ||
            (st = upload(e->d_ones, ones.data(), nr, e->s)) || (st = upload(e->d_nrebuild, zero.data(), nr, e->s)) ||
            (st = upload(e->d_rbany, zero.data(), 1, e->s)))
            return fail(st);
        if (e->ens()) {
            std::vector<int> sb(nr + 1, 0), cb(nr + 1, 0);
            for (int r = 0; r < nrep; ++r) {
                const RepMeta& m = e->meta[r];
                sb[r + 1] = sb[r] + m.nb1 + m.nb2 + m.nfl + m.nm;
                cb[r] = e->hreps[r].cube_off;
            }
            cb[nr] = (int)K;
            if (sb[nr] != (int)N) return fail(set_error(GC_ERR_INVALID, "EngineCreate: replica sizes"));
            // first and last replica of every 64-slot chunk, then of every 64-cube chunk
            const int nsc = chunks64((int)N), ncc = chunks64((int)K);
            std::vector<int2> cr((size_t)nsc + ncc);
            auto rep_at = [](const std::vector<int>& b, int x) {
                return (int)(std::upper_bound(b.begin(), b.end(), x) - b.begin()) - 1;
            };
            for (int c = 0; c < nsc; ++c)
                cr[c] = make_int2(rep_at(sb, 64 * c), rep_at(sb, std::min(64 * c + 63, (int)N - 1)));
            for (int c = 0; c < ncc; ++c)
                cr[(size_t)nsc + c] = make_int2(rep_at(cb, 64 * c), rep_at(cb, std::min(64 * c + 63, (int)K - 1)));
            const size_t nact = act_bytes((int)N) + act_bytes((int)K);
            if ((st = e->alloc(&e->d_slotbase, nr + 1)) || (st = e->alloc(&e->d_cubebase, nr + 1)) ||
                (st = e->alloc(&e->d_chunkrep, cr.size())) || (st = e->alloc(&e->d_gact, nact)) ||
                (st = hip_check(hipMemset(e->d_gact, 0, nact), "engine chunk activity")) ||
                (st = upload(e->d_chunkrep, cr.data(), cr.size(), e->s)) ||
                (st = upload(e->d_slotbase, sb.data(), nr + 1, e->s)) ||
                (st = upload(e->d_cubebase, cb.data(), nr + 1, e->s)))
                return fail(st);
        }
    }
    hipStream_t s = e->s;
    std::vector<float4> vout_id(N);
    std::vector<tfloat3> vh_id(N), f_id(N, tfloat3{0.f, 0.f, 0.f});
    std::vector<double> skin_pos;            // skin mode: the given positions (id order, SoA)
    std::vector<unsigned char> skin_code;    // and their offsets to the build's frame
    {
        // Particle::SetParticle (Particle.cu:485-546): rho = DenInit, F = 0
        std::vector<double> pos(3 * N);
        std::vector<tint3> nt(N, tint3{0, 0, 0});
        std::vector<int> idn(N), zero(K + 1, 0), rep_of(N);
        for (int r = 0; r < nrep; ++r) {
            const gc_case_desc& d = descs[r];
            const RepMeta& m = e->meta[r];
            const GcRep& R = e->hreps[r];
            const int lstart[5] = {0, m.nb1, m.nb1 + m.nb2, m.nb1 + m.nb2 + m.nfl,
                                   m.nb1 + m.nb2 + m.nfl + m.nm};
            for (int t = 0; t < 4; ++t)
                for (int l = lstart[t]; l < lstart[t + 1]; ++l) {
                    size_t g = (size_t)(R.id_off[t] + (l - lstart[t]));
                    pos[g] = d.coor[l].x; pos[N + g] = d.coor[l].y; pos[2 * N + g] = d.coor[l].z;
                    tfloat3 v = d.vel ? d.vel[l] : tfloat3{0.f, 0.f, 0.f};
                    if (t == 0) v = tfloat3{0.f, 0.f, 0.f};
                    vout_id[g] = make_float4(v.x, v.y, v.z, d.consts.DenInit);
                    vh_id[g] = t < 2 ? tfloat3{0.f, 0.f, 0.f} : v;
                    rep_of[g] = r;
                }
        }
        for (size_t i = 0; i < N; ++i) idn[i] = (int)i;
        if (e->skin_mode()) {   // the build runs on the wrapped positions (skin_frame)
            skin_pos = pos;
            skin_frame(e, rep_of, skin_pos, pos, skin_code);
        }
        unsigned long long ctr = e->ext.noise_counter;
        if ((st = upload(e->d_reps, e->hreps.data(), (size_t)nrep, s)) ||
            (st = upload(e->d_bbox, e->hreps[0].bbox, 6, s)) ||
            (st = upload(e->d_rep_of, rep_of.data(), N, s)) ||
            (st = upload(e->d_pos64[0], pos.data(), 3 * N, s)) ||
            (st = upload(e->d_vout, vout_id.data(), N, s)) || (st = upload(e->d_force, f_id.data(), N, s)) ||
            (st = upload(e->d_nt[0], nt.data(), N, s)) || (st = upload(e->d_sp[0], idn.data(), N, s)) ||
            (st = upload(e->d_curpos, idn.data(), N, s)) || (st = upload(e->d_to, idn.data(), N, s)) ||
            (st = upload(e->d_cnt, zero.data(), K + 1, s)) ||
            (st = upload(e->d_ctr, &ctr, 1, s)))
            return fail(st);
    }
    // first neighbour build (Particle::ConCubeList at init, Particle.cu:570)
    if ((st = launch_build(e, 0))) return fail(st);
    if (e->skin_mode() && e->cur == 1) {   // buffer 0 stays current in skin mode
        k3_copyback<<<gated_grid(blocks_for(e->N, BS)), BS, 0, s>>>(e->N, e->d_ones, e->d_pos64[1], e->d_pos64[0],
            e->d_nt[1], e->d_nt[0], e->d_sp[1], e->d_sp[0], nullptr, nullptr, nullptr);
        if ((st = hip_check(hipGetLastError(), "engine skin copy"))) return fail(st);
        e->cur = 0;
    }
    if ((st = reorder_carried(e, vout_id, f_id, vh_id))) return fail(st);
    if ((st = wall_tables(e))) return fail(st);
    if ((st = skin_init(e))) return fail(st);
    if (e->skin_mode() && (st = skin_restore_stored(e, skin_pos, skin_code))) return fail(st);

    if (e->ncell > 0) {
        const size_t T = (size_t)e->ntri, E = (size_t)e->nedge, C = (size_t)e->ncell, M = (size_t)e->nm;
        if ((st = e->alloc(&e->d_tri, T)) || (st = e->alloc(&e->d_edge, E)) ||
            (st = e->alloc(&e->d_celltab, M)) || (st = e->alloc(&e->d_betri, C)) ||
            (st = e->alloc(&e->d_mu0, E)) || (st = e->alloc(&e->d_kc, E)) ||
            (st = e->alloc(&e->d_kd, T)) || (st = e->alloc(&e->d_ka, T)) || (st = e->alloc(&e->d_kv, T)) ||
            (st = e->alloc(&e->d_l0, E)) || (st = e->alloc(&e->d_lmax, E)) || (st = e->alloc(&e->d_kb, E)) ||
            (st = e->alloc(&e->d_mexp, E)) || (st = e->alloc(&e->d_p, E)) || (st = e->alloc(&e->d_kp, E)) ||
            (st = e->alloc(&e->d_th0, E)) || (st = e->alloc(&e->d_cth0, E)) || (st = e->alloc(&e->d_sth0, E)) ||
            (st = e->alloc(&e->d_trinor, T)) ||
            (st = e->alloc(&e->d_tricen, T)) || (st = e->alloc(&e->d_triarea, T)) ||
            (st = e->alloc(&e->d_reftriarea, T)) || (st = e->alloc(&e->d_refarea_para, T)) ||
            (st = e->alloc(&e->d_refvol_para, T)) || (st = e->alloc(&e->d_area_para, T)) ||
            (st = e->alloc(&e->d_vol_para, T)) || (st = e->alloc(&e->d_area, C)) ||
            (st = e->alloc(&e->d_vol, C)) || (st = e->alloc(&e->d_pararea, M)) ||
            (st = e->alloc(&e->d_defforce, M)) || (st = e->alloc(&e->d_aggforce, M)) ||
            (st = e->alloc(&e->d_ef, 4 * E)) || (st = e->alloc(&e->d_tf, 3 * T)) ||
            (st = e->alloc(&e->d_mpos, 3 * M)))
            return fail(st);
        // global membrane arrays: 1-based global vertex ids, global triangle ids
        std::vector<tint3> tri(T);
        std::vector<int6> edge(E);
        std::vector<int> celltab(M), cell_rep(C);
        std::vector<tint2> betri(C);
        std::vector<float> mu(E), kc(E), kd(T), ka(T), kv(T);
        for (int r = 0; r < nrep; ++r) {
            const gc_case_desc& d = descs[r];
            const RepMeta& m = e->meta[r];
            const int vo = m.vtx_off, to = m.tri_off;
            for (int t = 0; t < m.ntri; ++t) {
                tri[to + t] = tint3{d.tri[t].x + vo, d.tri[t].y + vo, d.tri[t].z + vo};
                kd[to + t] = d.hlarea_para[t]; ka[to + t] = d.hgarea_para[t]; kv[to + t] = d.hvol_para[t];
            }
            for (int q = 0; q < m.nedge; ++q) {
                const int6& E0 = d.edge[q];
                edge[m.edge_off + q] = int6{E0.x + vo, E0.y + vo, E0.z + vo, E0.k + vo, E0.v + to, E0.w + to};
                mu[m.edge_off + q] = d.shearmod_para[q]; kc[m.edge_off + q] = d.bendingmod_para[q];
            }
            for (int v = 0; v < m.nm; ++v) celltab[vo + v] = d.celltab[v] + m.cell_off;
            for (int cc = 0; cc < m.ncell; ++cc) {
                betri[m.cell_off + cc] = tint2{d.beginendtri[cc].x + to, d.beginendtri[cc].y + to};
                cell_rep[m.cell_off + cc] = r;
            }
        }
        std::vector<int> es, ee, ts, te2;
        if ((st = build_incidence(e->nm, e->nedge, edge.data(), e->ntri, tri.data(), es, ee, ts, te2)))
            return fail(st);
        std::vector<int4> cellr(C);
        e->cell_fused = cell_ranges(e->nm, e->nedge, e->ntri, e->ncell, celltab.data(), betri.data(),
                                    edge.data(), tri.data(), cellr.data());
        if (e->cell_fused &&
            ((st = e->alloc(&e->d_cellr, C)) || (st = upload(e->d_cellr, cellr.data(), C, s))))
            return fail(st);
        if ((st = e->alloc(&e->d_es, es.size())) || (st = e->alloc(&e->d_ee, ee.size())) ||
            (st = e->alloc(&e->d_ts, ts.size())) || (st = e->alloc(&e->d_te, te2.size())))
            return fail(st);
        std::vector<tfloat3> zero(M, tfloat3{0.f, 0.f, 0.f});
        if ((st = upload(e->d_tri, tri.data(), T, s)) || (st = upload(e->d_edge, edge.data(), E, s)) ||
            (st = upload(e->d_celltab, celltab.data(), M, s)) || (st = upload(e->d_betri, betri.data(), C, s)) ||
            (st = upload(e->d_cell_rep, cell_rep.data(), C, s)) ||
            (st = upload(e->d_mu0, mu.data(), E, s)) || (st = upload(e->d_kc, kc.data(), E, s)) ||
            (st = upload(e->d_kd, kd.data(), T, s)) || (st = upload(e->d_ka, ka.data(), T, s)) ||
            (st = upload(e->d_kv, kv.data(), T, s)) || (st = upload(e->d_es, es.data(), es.size(), s)) ||
            (st = upload(e->d_ee, ee.data(), ee.size(), s)) || (st = upload(e->d_ts, ts.data(), ts.size(), s)) ||
            (st = upload(e->d_te, te2.data(), te2.size(), s)) || (st = upload(e->d_defforce, zero.data(), M, s)) ||
            (st = upload(e->d_aggforce, zero.data(), M, s)))
            return fail(st);
        const int c = e->cur;
        // reference geometry and mechanics (Particle::SetPar, Particle.cu:476-479)
        k_geo_cell<<<e->ncell, GC_GEO_BS, 0, s>>>(e->hreps[0].P, -1L, e->ens() ? e->d_reps : nullptr,
            e->ens() ? e->d_cell_rep : nullptr, e->d_bbox, e->d_curpos, e->d_tri,
            e->d_betri, e->d_nt[c], e->d_coor, e->d_trinor, e->d_triarea, e->d_reftriarea,
            e->d_tricen, e->d_refarea_para, e->d_refvol_para, e->d_area_para, e->d_vol_para,
            e->d_area, e->d_vol);
        k_par_area<<<blocks_for(e->nm, BS), BS, 0, s>>>(e->nm, e->d_ts, e->d_te, e->d_triarea, e->d_pararea);
        for (int r = 0; r < nrep; ++r) {  // per replica: its WLC-POW constants and kT
            const RepMeta& m = e->meta[r];
            if (m.nedge == 0) continue;
            const int o = m.edge_off;
            k_mech_para<<<blocks_for(m.nedge, BS), BS, 0, s>>>(e->hreps[r].P, m.nedge, e->d_edge + o,
                e->d_curpos, e->d_coor, e->d_l0 + o, e->d_lmax + o, e->d_kc + o, e->d_mu0 + o,
                e->d_kb + o, e->d_mexp + o, e->d_p + o, e->d_kp + o, e->d_th0 + o);
        }
        k_sincos<<<blocks_for(E, BS), BS, 0, s>>>(E, e->d_th0, e->d_cth0, e->d_sth0);
        if ((st = hip_check(hipGetLastError(), "engine membrane init"))) return fail(st);
    }
    if ((st = hip_check(hipStreamSynchronize(s), "engine init sync"))) return fail(st);
    e->step = 0;
    return e;
}

}  // namespace

extern "C" gc_engine* gc_EngineCreate(const gc_case_desc* d, int* status)
{
    return create(d, 1, status);
}

extern "C" gc_engine* gc_EnsembleCreate(const gc_case_desc* descs, int nrep, int* status)
{
    return create(descs, nrep, status);
}

extern "C" void gc_EngineDestroy(gc_engine* e) { destroy(e); }

extern "C" int gc_EngineStep(gc_engine* e, int nsteps, int driveon)
{
    if (!e || nsteps < 0) return set_error(GC_ERR_INVALID, "EngineStep: bad arguments");
    for (int k = 0; k < nsteps; ++k) {
        int st = launch_step(e, driveon, nullptr);
        if (st) return st;
    }
    return GC_OK;
}

extern "C" int gc_EngineStepGraph(gc_engine* e, int nsteps, int driveon)
{
    if (!e || nsteps < 0) return set_error(GC_ERR_INVALID, "EngineStepGraph: bad arguments");
    if (nsteps == 0) return GC_OK;
    // membrane substepping replays from the graph when it reads nothing back
    // on the host: no sample recording, and the surrogate's fallback decided
    // on the device (the per-cell membrane kernel)
    if (e->ms && !gc::ms_graph_ok(e)) return gc_EngineStep(e, nsteps, driveon);
    // A step alternates the double buffers, so one graph covers two steps
    // starting from buffer 0.
    if (e->cur != 0) {
        int st = launch_step(e, driveon, nullptr);
        if (st) return st;
        if (--nsteps == 0) return GC_OK;
    }
    if (!e->gexec || e->g_driveon != driveon) {
        if (e->gexec) { (void)hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
        hipGraph_t g;
        int st = hip_check(hipStreamBeginCapture(e->s, hipStreamCaptureModeThreadLocal), "begin capture");
        if (st) return st;
        long step0 = e->step;
        int st1 = launch_step(e, driveon, nullptr);
        int st2 = st1 ? st1 : launch_step(e, driveon, nullptr);
        hipError_t ce = hipStreamEndCapture(e->s, &g);
        e->step = step0;
        if (st2) return st2;
        if ((st = hip_check(ce, "end capture"))) return st;
        st = hip_check(hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0), "graph instantiate");
        (void)hipGraphDestroy(g);
        if (st) return st;
        e->g_driveon = driveon;
    }
    int pairs = nsteps / 2;
    for (int k = 0; k < pairs; ++k) {
        int st = hip_check(hipGraphLaunch(e->gexec, e->s), "graph launch");
        if (st) return st;
        e->step += 2;
    }
    e->walls_dirty = e->any_walls;
    if (nsteps & 1) return launch_step(e, driveon, nullptr);
    return GC_OK;
}

// Host-side check of the slot maps the re-sort produces (sp: slot -> id,
// CurPos: membrane vertex id -> slot, act: the moving particles' slots).
extern "C" int gc_CheckSlotMaps(int n, int nf, int nc, int nmov, const int* sp, const int* curpos,
                                const int* act)
{
    if (n < 0 || !sp || nf < 0 || nc < nf || nc > n || nmov < 0 || nmov > n || (nmov > 0 && !act) ||
        (nc < n && !curpos))
        return set_error(GC_ERR_INVALID, "CheckSlotMaps: bad arguments");
    char msg[160];
    auto bad = [&](const char* fmt, int a, int b, int c) {
        snprintf(msg, sizeof(msg), fmt, a, b, c);
        return set_error(GC_ERR_INVALID, std::string("CheckSlotMaps: ") + msg);
    };
    std::vector<unsigned char> seen((size_t)n, 0);
    int movers = 0;
    for (int q = 0; q < n; ++q) {
        const int id = sp[q];
        if (id < 0 || id >= n) return bad("slot %d holds id %d, outside [0, %d)", q, id, n);
        if (seen[id]++) return bad("id %d is in two slots (the second is %d)%.0d", id, q, 0);
        movers += id >= nf;
    }
    for (int id = nc; id < n; ++id) {
        const int q = curpos[id];
        if (q < 0 || q >= n || sp[q] != id)
            return bad("CurPos of membrane vertex %d is %d, which holds id %d", id, q,
                       q >= 0 && q < n ? sp[q] : -1);
    }
    if (nmov == n && movers < n) {   // moving walls: the force pass visits every slot
        for (int t = 0; t < n; ++t)
            if (act[t] != t) return bad("act[%d] = %d (every slot expected in order)%.0d", t, act[t], 0);
        return GC_OK;
    }
    if (nmov != movers) return bad("act holds %d slots for %d moving particles%.0d", nmov, movers, 0);
    int prev = -1;
    for (int t = 0; t < nmov; ++t) {
        const int q = act[t];
        if (q <= prev || q >= n) return bad("act[%d] = %d after %d (slots in order expected)", t, q, prev);
        if (sp[q] < nf) return bad("act[%d] = slot %d, which holds wall %d", t, q, sp[q]);
        prev = q;
    }
    return GC_OK;
}

namespace {

int check_slot_maps(gc_engine* e)
{
    const int N = e->N;
    std::vector<int> sp(N), cp(N), act((size_t)e->nmov);
    int st = hip_check(hipStreamSynchronize(e->s), "slot maps sync");
    if (!st) st = hip_check(hipMemcpy(sp.data(), e->d_sp[e->cur], sizeof(int) * N, hipMemcpyDeviceToHost), "slot maps D2H");
    if (!st) st = hip_check(hipMemcpy(cp.data(), e->d_curpos, sizeof(int) * N, hipMemcpyDeviceToHost), "slot maps D2H");
    if (!st && e->nmov > 0)
        st = hip_check(hipMemcpy(act.data(), e->d_act, sizeof(int) * e->nmov, hipMemcpyDeviceToHost), "slot maps D2H");
    if (st) return st;
    const GcParams& P = e->hreps[0].P;
    return gc_CheckSlotMaps(N, P.nf, P.nc, e->nmov, sp.data(), cp.data(), act.data());
}

// with GPUCELL_CHECK_SLOTS=1 in the environment at creation: after every
// re-sort of an eager step (not while a graph is being captured)
int debug_check_slot_maps(gc_engine* e)
{
    if (!e->check_slots) return GC_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    int st = hip_check(hipStreamIsCapturing(e->s, &cs), "capture status");
    if (st || cs != hipStreamCaptureStatusNone) return st;
    return check_slot_maps(e);
}

}  // namespace

extern "C" int gc_EngineCheckSlotMaps(gc_engine* e)
{
    if (!e) return set_error(GC_ERR_INVALID, "EngineCheckSlotMaps: null");
    return check_slot_maps(e);
}

extern "C" int gc_EngineInfo(const gc_engine* e, gc_engine_info* info)
{
    if (!e || !info) return set_error(GC_ERR_INVALID, "EngineInfo: null");
    memset(info, 0, sizeof(*info));
    info->step = e->step;
    unsigned long long c = 0;
    int st = hip_check(hipStreamSynchronize(e->s), "EngineInfo sync");
    if (st) return st;
    if ((st = hip_check(hipMemcpy(&c, e->d_ctr, sizeof(c), hipMemcpyDeviceToHost), "EngineInfo D2H")))
        return st;
    info->noise_counter = c;
    if ((st = hip_check(hipMemcpy(&info->fault, e->d_fault, sizeof(int), hipMemcpyDeviceToHost), "EngineInfo D2H")) ||
        (st = hip_check(hipMemcpy(&info->nl_overflow, e->d_nlover, sizeof(int), hipMemcpyDeviceToHost),
                        "EngineInfo D2H")))
        return st;
    if (e->skin_mode()) {
        std::vector<int> nrb((size_t)e->nrep);
        if ((st = hip_check(hipMemcpy(nrb.data(), e->d_nrebuild, sizeof(int) * e->nrep, hipMemcpyDeviceToHost),
                            "EngineInfo D2H")))
            return st;
        for (int v : nrb) info->nrebuild += v;
    }
    info->skin = e->skin;
    info->nTotalPar = e->N; info->nCellPar = e->nm; info->nCell = e->ncell; info->sizecube = e->K;
    for (int d = 0; d < 3; ++d) {
        info->ncube[d] = e->hreps[0].nc[d]; info->outerregion[d] = e->hreps[0].L[d];
        info->invwid[d] = e->hreps[0].invw[d];
    }
    return GC_OK;
}

namespace {

// GC_ERR_NONFINITE while a kernel has raised a fault bit (after a sync).
int fault_status(gc_engine* e)
{
    int f = 0;
    int st = hip_check(hipStreamSynchronize(e->s), "fault sync");
    if (!st) st = hip_check(hipMemcpy(&f, e->d_fault, sizeof(int), hipMemcpyDeviceToHost), "fault D2H");
    if (st) return st;
    if (f & GC_FAULT_WLC)
        return set_error(GC_ERR_NONFINITE, "a membrane edge reached its WLC maximum length l_max "
                                           "(singular spring force); the state is not physical");
    if (f & GC_FAULT_NONFINITE)
        return set_error(GC_ERR_NONFINITE, "non-finite position or force in the engine state");
    return GC_OK;
}

// The whole state in global id order.
struct Snapshot {
    std::vector<int> sp;
    std::vector<double> pos;
    std::vector<float4> vo;
    std::vector<tfloat3> vh, f, co;
    std::vector<tint3> nt;
};

int snapshot(gc_engine* e, Snapshot& S)
{
    const int N = e->N, c = e->cur;
    int st = finish_walls(e);
    if (st) return st;
    if ((st = hip_check(hipStreamSynchronize(e->s), "GetState sync"))) return st;
    std::vector<int> sp(N);
    std::vector<double> pos(3 * (size_t)N);
    std::vector<float4> vo(N);
    std::vector<tfloat3> vh(N), f(N), co(N);
    std::vector<tint3> nt(N);
    auto dl = [&](void* h, const void* dv, size_t b) {
        return hip_check(hipMemcpy(h, dv, b, hipMemcpyDeviceToHost), "GetState D2H");
    };
    if ((st = dl(sp.data(), e->d_sp[c], sizeof(int) * N)) ||
        (st = dl(pos.data(), e->d_pos64[c], sizeof(double) * 3 * N)) ||
        (st = dl(vo.data(), e->d_vout, sizeof(float4) * N)) ||
        (st = dl(vh.data(), e->d_velhalf, sizeof(tfloat3) * N)) ||
        (st = dl(f.data(), e->d_force, sizeof(tfloat3) * N)) ||
        (st = dl(co.data(), e->d_coor, sizeof(tfloat3) * N)) ||
        (st = dl(nt.data(), e->d_nt[c], sizeof(tint3) * N)))
        return st;
    S.sp = sp;
    S.pos.assign(3 * (size_t)N, 0.0);
    S.vo.resize(N); S.vh.resize(N); S.f.resize(N); S.co.resize(N); S.nt.resize(N);
    for (int slot = 0; slot < N; ++slot) {
        int id = sp[slot];
        // skin mode keeps the FP32 copy of moving particles in posid (coor
        // only for membrane vertices): the same float of the FP64 position
        if (e->skin_mode())
            co[slot] = tfloat3{(float)pos[slot], (float)pos[(size_t)N + slot], (float)pos[2 * (size_t)N + slot]};
        for (int d = 0; d < 3; ++d) S.pos[3 * (size_t)id + d] = pos[(size_t)d * N + slot];
        // the force pass leaves wall slots alone: walls carry no force
        S.vo[id] = vo[slot]; S.vh[id] = vh[slot]; S.co[id] = co[slot];
        S.f[id] = id < e->hreps[0].P.nf ? tfloat3{0.f, 0.f, 0.f} : f[slot];
        S.nt[id] = nt[slot];
    }
    return GC_OK;
}

// global id of replica r's own id l
int global_id(const gc_engine* e, int r, int l)
{
    const RepMeta& m = e->meta[r];
    const GcRep& R = e->hreps[r];
    int t = l < m.nb1 ? 0 : (l < m.nb1 + m.nb2 ? 1 : (l < m.nb1 + m.nb2 + m.nfl ? 2 : 3));
    return R.id_off[t] + (l - R.loc_start[t]);
}

int export_state(gc_engine* e, int r, gc_host_state* o)
{
    Snapshot S;
    int st = snapshot(e, S);
    if (st) return st;
    const RepMeta& m = e->meta[r];
    const int n = (int)m.consts.nTotalPar;
    const float c1sq = m.consts.SoundSpeed1 * m.consts.SoundSpeed1;
    const float c2sq = m.consts.SoundSpeed2 * m.consts.SoundSpeed2;
    for (int l = 0; l < n; ++l) {
        int g = global_id(e, r, l);
        if (o->pos) for (int d = 0; d < 3; ++d) o->pos[3 * (size_t)l + d] = S.pos[3 * (size_t)g + d];
        if (o->coor) o->coor[l] = S.co[g];
        if (o->velrhop) o->velrhop[l] = tfloat4{S.vo[g].x, S.vo[g].y, S.vo[g].z, S.vo[g].w};
        if (o->velhalf) o->velhalf[l] = S.vh[g];
        if (o->force) o->force[l] = S.f[g];
        if (o->press) o->press[l] = (l < (int)m.consts.nCellStart ? c1sq : c2sq) * S.vo[g].w;
        if (o->nthrobound) o->nthrobound[l] = S.nt[g];
    }
    if (m.ncell > 0) {
        auto dl = [&](void* h, const void* dv, size_t b) {
            return hip_check(hipMemcpy(h, dv, b, hipMemcpyDeviceToHost), "GetState D2H");
        };
        if (o->area && (st = dl(o->area, e->d_area + m.cell_off, sizeof(float) * m.ncell))) return st;
        if (o->vol && (st = dl(o->vol, e->d_vol + m.cell_off, sizeof(float) * m.ncell))) return st;
        if (o->defforce && (st = dl(o->defforce, e->d_defforce + m.vtx_off, sizeof(tfloat3) * m.nm))) return st;
        if (o->aggforce && (st = dl(o->aggforce, e->d_aggforce + m.vtx_off, sizeof(tfloat3) * m.nm))) return st;
    }
    return fault_status(e);
}

}  // namespace

extern "C" int gc_EngineGetState(gc_engine* e, gc_host_state* o)
{
    if (!e || !o) return set_error(GC_ERR_INVALID, "EngineGetState: null");
    if (e->nrep != 1) return set_error(GC_ERR_INVALID, "EngineGetState: batched engine (use EnsembleGetState)");
    return export_state(e, 0, o);
}

extern "C" int gc_EnsembleGetState(gc_engine* e, int rep, gc_host_state* o)
{
    if (!e || !o || rep < 0 || rep >= e->nrep) return set_error(GC_ERR_INVALID, "EnsembleGetState: bad arguments");
    return export_state(e, rep, o);
}

namespace {

// Loads a state given in global id order (positions AoS) and re-sorts.
int restore_state(gc_engine* e, const std::vector<double>& pos, const std::vector<float4>& vo,
                  const std::vector<tfloat3>& f, const std::vector<tfloat3>& vh,
                  const std::vector<tint3>& nt, long step, unsigned long long ctr)
{
    const int N = e->N;
    if (e->gexec) { (void)hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
    std::vector<double> pos_soa(3 * (size_t)N);
    for (int i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) pos_soa[(size_t)d * N + i] = pos[3 * (size_t)i + d];
    std::vector<double> skin_pos;
    std::vector<unsigned char> skin_code;
    if (e->skin_mode()) {   // the build runs on the wrapped positions (skin_frame)
        skin_pos = pos_soa;
        skin_frame(e, {}, skin_pos, pos_soa, skin_code);
    }
    std::vector<int> idn(N);
    for (int i = 0; i < N; ++i) idn[i] = i;
    e->cur = 0;
    hipStream_t s = e->s;
    int st;
    const int zero = 0;
    if ((st = upload(e->d_pos64[0], pos_soa.data(), 3 * (size_t)N, s)) || (st = upload(e->d_vout, vo.data(), N, s)) ||
        (st = upload(e->d_force, f.data(), N, s)) || (st = upload(e->d_nt[0], nt.data(), N, s)) ||
        (st = upload(e->d_sp[0], idn.data(), N, s)) || (st = upload(e->d_curpos, idn.data(), N, s)) ||
        (st = upload(e->d_ctr, &ctr, 1, s)) || (st = upload(e->d_fault, &zero, 1, s)))
        return st;
    // the build's histogram starts from zero counts (a batch in skin mode keeps them)
    if (e->skin_mode() && e->ens() &&
        (st = hip_check(hipMemsetAsync(e->d_cnt, 0, sizeof(int) * ((size_t)e->K + 1), s), "reset counts")))
        return st;
    if ((st = launch_build(e, 0))) return st;
    if (e->skin_mode() && e->cur == 1) {   // buffer 0 stays current in skin mode
        k3_copyback<<<gated_grid(blocks_for(N, BS)), BS, 0, s>>>(N, e->d_ones, e->d_pos64[1], e->d_pos64[0], e->d_nt[1],
                                                     e->d_nt[0], e->d_sp[1], e->d_sp[0], nullptr, nullptr, nullptr);
        if ((st = hip_check(hipGetLastError(), "engine skin copy"))) return st;
        e->cur = 0;
    }
    if ((st = reorder_carried(e, vo, f, vh))) return st;
    if ((st = wall_tables(e))) return st;
    if ((st = skin_init(e))) return st;
    if (e->skin_mode() && (st = skin_restore_stored(e, skin_pos, skin_code))) return st;
    e->walls_dirty = false;
    e->step = step;
    if (e->ms && (st = ms_reset(e))) return st;
    return GC_OK;
}

}  // namespace

extern "C" int gc_EngineSetState(gc_engine* e, const gc_host_state* in, long step,
                                 uint64_t noise_counter)
{
    if (!e || !in) return set_error(GC_ERR_INVALID, "EngineSetState: null");
    if (e->nrep != 1) return set_error(GC_ERR_INVALID, "EngineSetState: batched engine");
    const int N = e->N;
    int st = hip_check(hipStreamSynchronize(e->s), "SetState sync");
    if (st) return st;
    Snapshot S;
    if ((st = snapshot(e, S))) return st;
    std::vector<double> pos = S.pos;
    if (in->pos) memcpy(pos.data(), in->pos, sizeof(double) * 3 * N);
    std::vector<float4> vo = S.vo;
    if (in->velrhop)
        for (int i = 0; i < N; ++i)
            vo[i] = make_float4(in->velrhop[i].x, in->velrhop[i].y, in->velrhop[i].z, in->velrhop[i].w);
    std::vector<tfloat3> vh = S.vh, f = S.f;
    std::vector<tint3> nt = S.nt;
    if (in->velhalf) memcpy(vh.data(), in->velhalf, sizeof(tfloat3) * N);
    if (in->force) memcpy(f.data(), in->force, sizeof(tfloat3) * N);
    if (in->nthrobound) memcpy(nt.data(), in->nthrobound, sizeof(tint3) * N);
    return restore_state(e, pos, vo, f, vh, nt, step, noise_counter);
}

extern "C" int gc_EngineGetCells(gc_engine* e, gc_host_cells* o)
{
    if (!e || !o) return set_error(GC_ERR_INVALID, "EngineGetCells: null");
    int st = hip_check(hipStreamSynchronize(e->s), "GetCells sync");
    if (st) return st;
    const int N = e->N, c = e->cur, K = e->K;
    auto dl = [&](void* h, const void* dv, size_t b) {
        return h ? hip_check(hipMemcpy(h, dv, b, hipMemcpyDeviceToHost), "GetCells D2H") : GC_OK;
    };
    std::vector<int> key(N), sp(N);
    if ((st = dl(sp.data(), e->d_sp[c], sizeof(int) * N)) ||
        (st = dl(key.data(), e->d_key, sizeof(int) * N)))
        return st;
    // The reference's cell list (keys x + nx (y + ny z)): the engine groups
    // its rows into tiles, so its slots are listed here by reference key (a
    // counting sort; a cube's slots stay in id order).
    std::vector<int> ref(N), start(K + 1, 0), order(N);
    std::vector<tint3> cc(N);
    for (int s = 0, r = 0; s < N; ++s) {  // slots of a replica are contiguous
        while (r + 1 < e->nrep && key[s] >= e->hreps[r + 1].cube_off) ++r;
        const GcRep& R = e->hreps[r];
        if (key[s] < R.cube_off || key[s] >= K) return set_error(GC_ERR_GRID, "EngineGetCells: key out of range");
        int cx, cy, cz;
        cube_coords(R, key[s], cx, cy, cz);
        cc[s] = tint3{cx, cy, cz};
        ref[s] = R.cube_off + cx + R.nc[0] * (cy + R.nc[1] * cz);
        start[ref[s] + 1] += 1;
    }
    for (int k = 0; k < K; ++k) start[k + 1] += start[k];
    {
        std::vector<int> pos(start.begin(), start.end() - 1);
        for (int s = 0; s < N; ++s) order[pos[ref[s]]++] = s;
    }
    for (int i = 0; i < N; ++i) {
        const int s = order[i];
        if (o->sortpart) o->sortpart[i] = sp[s];
        if (o->cubepart) o->cubepart[i] = ref[s];
        if (o->cc) o->cc[i] = cc[s];
        // the step keeps CurPos on the device for membrane vertices only;
        // the whole map is the inverse of SortPart
        if (o->curpos) o->curpos[sp[s]] = i;
    }
    if (o->beginendcube)
        for (int k = 0; k < K; ++k) o->beginendcube[k] = tint2{start[k], start[k + 1]};
    return GC_OK;
}

extern "C" int gc_EngineProfile(gc_engine* e, int nsteps, int driveon, gc_engine_profile* out)
{
    if (!e || !out || nsteps <= 0) return set_error(GC_ERR_INVALID, "EngineProfile: bad arguments");
    memset(out, 0, sizeof(*out));
    Prof pr;
    for (int k = 0; k < 32; ++k) (void)hipEventCreate(&pr.ev[k]);
    std::vector<double> acc(32, 0.0);
    int nmarks = 0;
    const char* names[32] = {nullptr};
    int st = GC_OK;
    for (int it = 0; it < nsteps && !st; ++it) {
        pr.n = 0;
        st = launch_step(e, driveon, &pr);
        if (st) break;
        st = hip_check(hipEventSynchronize(pr.ev[pr.n - 1]), "profile sync");
        nmarks = pr.n;
        for (int k = 1; k < pr.n; ++k) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, pr.ev[k - 1], pr.ev[k]);
            acc[k] += ms;
            names[k] = pr.name[k];
        }
    }
    for (int k = 0; k < 32; ++k) (void)hipEventDestroy(pr.ev[k]);
    if (st) return st;
    int m = 0;
    for (int k = 1; k < nmarks && m < 16; ++k, ++m) {
        strncpy(out->names[m], names[k], 31);
        out->us[m] = (float)(acc[k] * 1000.0 / nsteps);
    }
    out->nkernels = m;
    return GC_OK;
}

namespace {

const char* const TUNE_PHASE[TUNE_N] = {"integrate+key", "density+wall velocity", "sdpd force",
                                        "membrane force+corrector"};
// a second phase timed with the first (the per-cell kernel also does the geometry)
const char* const TUNE_PHASE2[TUNE_N] = {nullptr, nullptr, nullptr, "membrane geometry"};
const char* const TUNE_FAMILY[TUNE_N] = {"integrate", "density", "sdpd force", "membrane force"};

void tune_shape(int f, int v, char* out, size_t n)
{
    static const int dbs[12] = {64, 64, 64, 64, 128, 128, 128, 128, 256, 256, 256, 256};
    static const int dst[12] = {0, 128, 192, 320, 0, 128, 192, 320, 0, 128, 192, 320};
    if (f == TUNE_DENSITY) snprintf(out, n, "block %d, stage %d", dbs[v], dst[v]);
    else if (f == TUNE_MEMBRANE && v < 3) snprintf(out, n, "3 passes, group %d", 64 << v);
    else if (f == TUNE_MEMBRANE) snprintf(out, n, "per cell, %d threads", GC_CELL_BS);
    else snprintf(out, n, "block %d", 64 << v);
}

// Average time (us) of one phase (plus name2 if given) over nsteps eager steps.
int phase_us(gc_engine* e, int nsteps, int driveon, const char* name, const char* name2, float* us)
{
    Prof pr;
    for (int k = 0; k < 32; ++k) (void)hipEventCreate(&pr.ev[k]);
    double acc = 0.0;
    int st = GC_OK;
    for (int it = 0; it < nsteps && !st; ++it) {
        pr.n = 0;
        if ((st = launch_step(e, driveon, &pr))) break;
        st = hip_check(hipEventSynchronize(pr.ev[pr.n - 1]), "autotune sync");
        for (int k = 1; k < pr.n && !st; ++k)
            if (!strcmp(pr.name[k], name) || (name2 && !strcmp(pr.name[k], name2))) {
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, pr.ev[k - 1], pr.ev[k]);
                acc += ms;
            }
    }
    for (int k = 0; k < 32; ++k) (void)hipEventDestroy(pr.ev[k]);
    *us = (float)(acc * 1000.0 / (nsteps > 0 ? nsteps : 1));
    return st;
}

}  // namespace

extern "C" int gc_EngineAutotune(gc_engine* e, int steps, int driveon, gc_tune_report* out)
{
    if (!e || steps < 1) return set_error(GC_ERR_INVALID, "EngineAutotune: bad arguments");
    if (e->ms) return set_error(GC_ERR_INVALID, "EngineAutotune: membrane substepping active");
    int st = hip_check(hipStreamSynchronize(e->s), "Autotune sync");
    if (st) return st;
    // the state is restored afterwards: tuning leaves the trajectory untouched
    Snapshot S;
    if ((st = snapshot(e, S))) return st;
    const long step0 = e->step;
    unsigned long long ctr0 = 0;
    if ((st = hip_check(hipMemcpy(&ctr0, e->d_ctr, sizeof(ctr0), hipMemcpyDeviceToHost), "Autotune ctr")))
        return st;
    if (out) memset(out, 0, sizeof(*out));
    int row = 0;
    for (int f = 0; f < TUNE_N && !st; ++f) {
        if (f == TUNE_MEMBRANE && e->ncell == 0) continue;
        float best = 3.0e38f;
        int bestv = e->tune.v[f];
        const int row0 = row;
        for (int v = 0; v < TUNE_NVARIANT[f] && !st; ++v) {
            if (f == TUNE_MEMBRANE && v == 3 && !e->cell_fused) continue;   // cells too large
            // skin mode: the stage only shapes the list build, which the timed
            // steps rarely run; the density pass itself has no stage
            if (f == TUNE_DENSITY && e->skin_mode() && v % 4 != 2) continue;
            e->tune.v[f] = v;
            float us = 0.f;
            if ((st = launch_step(e, driveon, nullptr))) break;   // warm-up (code load)
            const char* ph = f == TUNE_INTEGRATE && e->skin_mode() ? "advance+predict" : TUNE_PHASE[f];
            if ((st = phase_us(e, steps, driveon, ph, TUNE_PHASE2[f], &us))) break;
            if (us < best) { best = us; bestv = v; }
            if (out && row < 64) {
                snprintf(out->family[row], sizeof(out->family[row]), "%s", TUNE_FAMILY[f]);
                out->variant[row] = v;
                tune_shape(f, v, out->shape[row], sizeof(out->shape[row]));
                out->us[row] = us;
                ++row;
            }
        }
        e->tune.v[f] = bestv;
        if (out)
            for (int r = row0; r < row; ++r) out->chosen[r] = out->variant[r] == bestv;
    }
    if (out) out->n = row;
    int st2 = restore_state(e, S.pos, S.vo, S.f, S.vh, S.nt, step0, ctr0);
    return st ? st : st2;
}

extern "C" int gc_EngineGetTuning(const gc_engine* e, int* v)
{
    if (!e || !v) return set_error(GC_ERR_INVALID, "EngineGetTuning: null");
    for (int f = 0; f < TUNE_N; ++f) v[f] = e->tune.v[f];
    return GC_OK;
}

extern "C" int gc_EngineSetTuning(gc_engine* e, const int* v)
{
    if (!e || !v) return set_error(GC_ERR_INVALID, "EngineSetTuning: null");
    for (int f = 0; f < TUNE_N; ++f)
        if (v[f] < 0 || v[f] >= TUNE_NVARIANT[f]) return set_error(GC_ERR_INVALID, "EngineSetTuning: variant out of range");
    int st = hip_check(hipStreamSynchronize(e->s), "SetTuning sync");
    if (st) return st;
    if (e->gexec) { (void)hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
    for (int f = 0; f < TUNE_N; ++f) e->tune.v[f] = v[f];
    return GC_OK;
}

extern "C" int gc_EngineGetTriangles(gc_engine* e, tfloat3* trinordir, tfloat3* tricenter)
{
    if (!e) return set_error(GC_ERR_INVALID, "EngineGetTriangles: null");
    if (e->ncell == 0) return GC_OK;
    int st = hip_check(hipStreamSynchronize(e->s), "GetTriangles sync");
    if (st) return st;
    if (trinordir && (st = hip_check(hipMemcpy(trinordir, e->d_trinor, sizeof(tfloat3) * e->ntri,
                                               hipMemcpyDeviceToHost), "GetTriangles nor")))
        return st;
    if (tricenter && (st = hip_check(hipMemcpy(tricenter, e->d_tricen, sizeof(tfloat3) * e->ntri,
                                               hipMemcpyDeviceToHost), "GetTriangles cen")))
        return st;
    return GC_OK;
}

extern "C" int gc_EngineRanges(const gc_engine* e, int* nbound1, int* nbound12)
{
    if (!e || !nbound1 || !nbound12) return set_error(GC_ERR_INVALID, "EngineRanges: null");
    *nbound1 = (int)e->consts.nBound2Start;
    *nbound12 = (int)e->consts.nFluidStart;
    return GC_OK;
}

extern "C" int gc_EnsembleCount(const gc_engine* e) { return e ? e->nrep : 0; }
