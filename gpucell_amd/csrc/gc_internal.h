// gc_internal.h -- internals shared by the HIP translation units of
// libgpucell_hip.so: the per-launch parameter block, device math (Lucy
// kernel, Philox-4x32-10, Box-Muller), the cube-range visitor of the
// neighbour loops, and host helpers (status, stream, scratch).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/gpucell/gpucell.h"

// Work-group size of the per-particle kernels: 4 wave64 per group.
#define GC_BS 256

// MI355X has 8 XCDs, each with its own L2; the dispatcher deals work-groups
// to XCDs round-robin.  xcd_block() renumbers them so that XCD x runs one
// contiguous run of logical blocks: neighbouring slots, and the neighbour
// cubes they read, then share an L2 instead of being fetched by all eight.
// Launch xcd_grid(nb) groups; groups whose logical id is >= nb return.
#define GC_NXCD 8
inline int xcd_grid(int nb) { return ((nb + GC_NXCD - 1) / GC_NXCD) * GC_NXCD; }
__device__ __forceinline__ int xcd_block(int nb)
{
    int per = (nb + GC_NXCD - 1) / GC_NXCD;
    return (int)(blockIdx.x % GC_NXCD) * per + (int)(blockIdx.x / GC_NXCD);
}

// ---------------------------------------------------------------------------
// Parameter block: Structconst + gc_ext_consts + derived constants, passed
// by value as a kernel argument (kernarg segment: scalar loads).
struct GcParams {
    int nb2, nf, nc, N, nm;     // id ranges (TypesDef.cuh:52-56)
    float h, h2, m, inv_m;
    float c1sq, c2sq, rho0, lam, fext, dt;
    float wnorm;                // Lucy W prefactor 105/(16 pi h^3)
    float fnorm;                // Lucy F prefactor 315/(4 pi h^5)
    float visc_a, visc_b;       // 5eta/3 - zeta, 5(zeta + eta/3)
    float sq_a, sq_ab;          // sqrt(a), sqrt(a + b)
    float two_kT_over_dt;       // 2 kT / dt
    float krep;                 // Bound1 repulsion coefficient
    float De, mcut2, r0, beta, kT;
    float force_split;
    float wall_split, wall_u_lo, wall_u_hi;   // moving Bound2 walls (x velocity below / above the split)
    float off[3];               // key padding offset per axis
    int per[3];
    int force_mode;
    int noise_on;
    float wlc_ratio, pow_index;
    uint32_t seed_lo, seed_hi;
    uint32_t ctr_lo, ctr_hi;    // noise counter of this evaluation
};

// Neighbour-loop view of the cube grid.
struct GcGridV {
    int nc[3];
    float L[3];                 // OuterRegion (period on periodic axes)
};

// ---------------------------------------------------------------------------
// Device helpers
__device__ __forceinline__ float lucy_w(float r, const GcParams& P)
{
    float q = r * (1.0f / P.h);
    float om = 1.0f - q;
    return q < 1.0f ? P.wnorm * (1.0f + 3.0f * q) * om * om * om : 0.0f;
}

// Velocity of a Bound2 wall particle at height y from the Lucy-weighted
// sums of its moving neighbours' velocities (sw, svx, svy, svz): the no-slip
// image 2 u_w - sum(W v) / sum(W) about the wall's own velocity u_w (along
// x: wall_u_lo below wall_split, wall_u_hi above); u_w with no moving
// neighbour.  For walls at rest this is -sum(W v) / sum(W).
__device__ __forceinline__ float3 wall_velocity(const GcParams& P, float y, float sw, float svx, float svy,
                                                float svz)
{
    const float uw = y >= P.wall_split ? P.wall_u_hi : P.wall_u_lo;
    const float s = sw > 0.f ? -1.0f / sw : 0.0f;
    float3 v = make_float3(svx * s, svy * s, svz * s);
    if (uw != 0.0f) v.x += sw > 0.f ? 2.0f * uw : uw;
    return v;
}

// F(r) = -W'(r)/r
__device__ __forceinline__ float lucy_f(float r, const GcParams& P)
{
    float q = r * (1.0f / P.h);
    float om = 1.0f - q;
    return q < 1.0f ? P.fnorm * om * om : 0.0f;
}

__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                              uint32_t c3, uint32_t k0, uint32_t k1,
                                              uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one v_mad_u64_u32 per product instead of a mul_lo + mul_hi pair
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ float u01f(uint32_t x)
{
    return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

// Three standard normals of the pair {ida, idb}: Philox counter
// (min id, max id, counter lo, counter hi), key (seed lo, seed hi),
// Box-Muller on (u0, u1) and (u2, u3).  The transcendentals are the
// hardware instructions: v_log_f32 (log2), v_sqrt_f32 and v_sin_f32 /
// v_cos_f32, whose argument is in revolutions, so cos(2 pi u) is one
// instruction.
__device__ __forceinline__ void pair_normals_k(uint32_t seed_lo, uint32_t seed_hi, uint32_t ctr_lo,
                                               uint32_t ctr_hi, uint32_t ida, uint32_t idb,
                                               float& n0, float& n1, float& n2)
{
    uint32_t lo = min(ida, idb), hi = max(ida, idb);
    uint32_t o[4];
    philox4x32_10(lo, hi, ctr_lo, ctr_hi, seed_lo, seed_hi, o);
    const float m2ln2 = -1.3862943611198906f;  // -2 ln 2
    float r0 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u01f(o[0])));
    float r1 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u01f(o[2])));
    float u1 = u01f(o[1]), u3 = u01f(o[3]);
    n0 = r0 * __builtin_amdgcn_cosf(u1);
    n1 = r0 * __builtin_amdgcn_sinf(u1);
    n2 = r1 * __builtin_amdgcn_cosf(u3);
}

__device__ __forceinline__ void pair_normals(const GcParams& P, uint32_t ida, uint32_t idb,
                                             float& n0, float& n1, float& n2)
{
    pair_normals_k(P.seed_lo, P.seed_hi, P.ctr_lo, P.ctr_hi, ida, idb, n0, n1, n2);
}

// Cube coordinate of one axis, FP32 with no contraction so the host oracle
// reproduces it bit for bit: floor(((x - lo) + off) * invw), clamped.
// x_i - (x_j + s) for a periodic image shift s in {-L, 0, +L} of a
// neighbour j, with a single rounding at the end.  The shift is applied to
// the coordinate next to the far face of the box (with lo = 0 it lies in
// [L/2, L], so subtracting L is exact by Sterbenz's lemma): s = -L means x_j
// is near L, s = +L means x_i is.  Forming x_j + L instead would round to
// the spacing of floats near L (7.6e-6 at L = 230), an error in the pair
// distance of the same size for every pair across a periodic face.
// s = 0 gives x_i - x_j.
__device__ __forceinline__ float pdiff(float xi, float xj, float s)
{
    return s > 0.0f ? (xi - s) - xj : xi - (xj + s);
}

__device__ __forceinline__ int cube_coord(float x, float lo, float off, float invw, int nc)
{
#pragma clang fp contract(off)
    float t = x - lo;
    float u = t + off;
    float f = u * invw;
    int k = (int)floorf(f);
    k = k < 0 ? 0 : k;
    k = k > nc - 1 ? nc - 1 : k;
    return k;
}

// Visits the neighbour cubes of (cx, cy, cz) as contiguous slot ranges:
// keys are x-fastest, so the three cubes of one (y, z) row that do not
// wrap form one range [begin(x-1), end(x+1)).  fn(jb, je, sx, sy, sz)
// receives the image shift to add to a neighbour's coordinates.
template <class Fn>
__device__ __forceinline__ void visit_ranges(const GcGridV& g, const int* per,
                                             const tint2* __restrict__ be, int cx, int cy,
                                             int cz, Fn&& fn)
{
    for (int dz = -1; dz <= 1; ++dz) {
        int z = cz + dz;
        float sz = 0.0f;
        if (z < 0) { if (!per[2]) continue; z += g.nc[2]; sz = -g.L[2]; }
        else if (z >= g.nc[2]) { if (!per[2]) continue; z -= g.nc[2]; sz = g.L[2]; }
        for (int dy = -1; dy <= 1; ++dy) {
            int y = cy + dy;
            float sy = 0.0f;
            if (y < 0) { if (!per[1]) continue; y += g.nc[1]; sy = -g.L[1]; }
            else if (y >= g.nc[1]) { if (!per[1]) continue; y -= g.nc[1]; sy = g.L[1]; }
            int row = (z * g.nc[1] + y) * g.nc[0];
            int xlo = cx - 1, xhi = cx + 1;
            if (xlo < 0) {
                if (per[0]) { tint2 r = be[row + g.nc[0] - 1]; fn(r.x, r.y, -g.L[0], sy, sz); }
                xlo = 0;
            }
            if (xhi >= g.nc[0]) {
                if (per[0]) { tint2 r = be[row]; fn(r.x, r.y, g.L[0], sy, sz); }
                xhi = g.nc[0] - 1;
            }
            fn(be[row + xlo].x, be[row + xhi].y, 0.0f, sy, sz);
        }
    }
}

// ---------------------------------------------------------------------------
// Host helpers (defined in gc_abi.hip)
namespace gc {

int set_error(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);
hipStream_t stream();
// Current parameter block; GC_ERR_NOCONSTS if gc_ConstsDataUp was not called.
int params(GcParams* out);
// Builds a parameter block from explicit blocks (engine use).
void make_params(const Structconst& c, const gc_ext_consts& ext, GcParams* out);
// Library-owned growable scratch buffer number `slot` of at least `bytes`.
int scratch(int slot, size_t bytes, void** ptr);
// Advances the global noise counter (one per SDPD force evaluation).
void advance_noise_counter();

inline int blocks_for(long n, int bs) { return (int)((n + bs - 1) / bs); }

// Counting sort of (key, value) by key, ties by ascending value.
// keys in [0, K); writes sorted keys/values into out_k/out_v (may alias the
// inputs only through the internal copies).  cnt must hold K+1 ints.
// scan_ws: workspace of scan_ws_ints(K) ints, or NULL for library scratch.
int counting_sort(int n, int K, const int* keys, const int* vals, int* out_k, int* out_v,
                  int* cnt, int* offs, int* rank, int* tmp_k, int* tmp_v, hipStream_t s,
                  int* scan_ws = nullptr);
// Exclusive scan of n ints (in != out), total in out[n].  ws: workspace of
// scan_ws_ints(n) ints, or NULL for library scratch.
int exclusive_scan(const int* in, int* out, int n, hipStream_t s, int* ws = nullptr);
size_t scan_ws_ints(int n);

}  // namespace gc

#define GC_LAUNCH_CHECK(what) \
    do { hipError_t _e = hipGetLastError(); if (_e != hipSuccess) return gc::hip_check(_e, what); } while (0)
