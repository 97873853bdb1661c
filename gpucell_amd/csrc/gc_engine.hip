// gc_engine.hip -- the engine: library-owned state and the step of
// Particle::UpdatePar (src/Particle.cu:1310-1329) as a fixed sequence of
// fused kernels on one stream, captured into a HIP graph for replay.
//
// Per step (reference launch numbers of UpdatePar in brackets):
//   k2_integrate   [1,2,3]  r += dt v + dt^2/2 F/m in FP64, periodic wrap,
//                           cube key, histogram rank (atomic)
//   scan                    cube offsets (exclusive scan of the histogram)
//   k2_scatter     [4]      (cube, id) pairs into cube segments
//   k2_sortgather  [4,5,6,8] rank within the cube by id -> final slot;
//                           gather of the state; Groot-Warren predictor
//   k2_density     [7,9]    number density (moving) and no-slip wall
//                           velocity (Bound2) in one neighbour pass
//   k2_force       [10-13,17] EOS on the fly, SDPD pair force, Bound1
//                           repulsion, body force; corrector for fluid
//   k_geo_cell     [14]     membrane geometry, FP64 cell area/volume
//   k2_agg         [16]     Morse between cells (when De != 0)
//   k_def_force    [15,17]  membrane forces, corrector for membrane
//   k2_wallvel     [18]     wall velocity from the corrected flow
//
// Layouts (slot order): posid float4 (x, y, z, id bits), velrho float4
// (predicted v, rho) read by neighbour loops, vout float4 (v(t+dt), rho) =
// the reference's Velrhop after the step, pos64 = x[N], y[N], z[N] doubles (accumulator),
// coor/force/velhalf in the TypesDef float3 layout.  BeginEndCube is the
// scan's offsets array: cube k holds slots [offs[k], offs[k+1]).
#include "gc_kernels.h"

#include <math.h>
#include <string.h>
#include <vector>

using namespace gc;

namespace {

constexpr int BS = GC_BS;
// Hits buffered per lane by the force kernel (LDS: FORCE_HMAX * BS ints).
constexpr int FORCE_HMAX = 32;

struct Grid2 {
    int nc[3];
    float L[3];
};

__device__ __forceinline__ void key_to_cc(int k, const Grid2& g, int& cx, int& cy, int& cz)
{
    cx = k % g.nc[0];
    int t = k / g.nc[0];
    cy = t % g.nc[1];
    cz = t / g.nc[1];
}

// Contiguous slot ranges of the 27 neighbour cubes (x-fastest keys: one
// range per non-wrapping (y, z) row), with the image shift of each range.
template <class Fn>
__device__ __forceinline__ void visit2(const Grid2& g, const int* per, const int* __restrict__ offs,
                                       int cx, int cy, int cz, Fn&& fn)
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
                if (per[0]) fn(offs[row + g.nc[0] - 1], offs[row + g.nc[0]], -g.L[0], sy, sz);
                xlo = 0;
            }
            if (xhi >= g.nc[0]) {
                if (per[0]) fn(offs[row], offs[row + 1], g.L[0], sy, sz);
                xhi = g.nc[0] - 1;
            }
            fn(offs[row + xlo], offs[row + xhi + 1], 0.0f, sy, sz);
        }
    }
}

__device__ __forceinline__ float ext_force_x(const GcParams& P, float y)
{
    float f = P.fext;
    return (P.force_mode == GC_FORCE_REVERSE_POISEUILLE && y >= P.force_split) ? -f : f;
}

// [1,2,3] integrate (FP64), wrap, key, histogram rank
__global__ __launch_bounds__(BS) void k2_integrate(int n, GcParams P, int move,
    double* __restrict__ pos64, const float4* __restrict__ vout, const tfloat3* __restrict__ force,
    const int* __restrict__ sp, tint3* __restrict__ nthro, const float* __restrict__ bbox,
    const float* __restrict__ invw, const int* __restrict__ ncube, int* __restrict__ key,
    int* __restrict__ rank, int* __restrict__ cnt)
{
    int p = xcd_block((n + BS - 1) / BS) * BS + threadIdx.x;
    if (p >= n) return;
    int id = sp[p];
    double x = pos64[p], y = pos64[n + p], z = pos64[2 * n + p];
    if (move && id >= P.nf) {
        double dt = P.dt, a = 0.5 * (double)P.dt * (double)P.dt * (double)P.inv_m;
        float4 v = vout[p];
        tfloat3 f = force[p];
        double u[3] = {x + (dt * v.x + a * f.x), y + (dt * v.y + a * f.y), z + (dt * v.z + a * f.z)};
        tint3 im = nthro[p];
        int img[3] = {im.x, im.y, im.z};
        bool wrapped = false;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (!P.per[d]) continue;
            double lo = bbox[2 * d], hi = bbox[2 * d + 1];
            double L = (double)(bbox[2 * d + 1] - bbox[2 * d]);
            if (u[d] < lo) { u[d] += L; img[d] -= 1; wrapped = true; }
            else if (u[d] >= hi) { u[d] -= L; img[d] += 1; wrapped = true; }
        }
        if (wrapped) nthro[p] = tint3{img[0], img[1], img[2]};
        x = u[0]; y = u[1]; z = u[2];
        pos64[p] = x; pos64[n + p] = y; pos64[2 * n + p] = z;
    }
    int nx = ncube[0], ny = ncube[1], nz = ncube[2];
    int cx = cube_coord((float)x, bbox[0], P.off[0], invw[0], nx);
    int cy = cube_coord((float)y, bbox[2], P.off[1], invw[1], ny);
    int cz = cube_coord((float)z, bbox[4], P.off[2], invw[2], nz);
    int k = cx + nx * (cy + ny * cz);
    key[p] = k;
    rank[p] = atomicAdd(&cnt[k], 1);
}

// [4] (cube, id) into cube segments; clears the histogram for the next step
// (every non-empty cube has exactly one rank-0 member).
__global__ __launch_bounds__(BS) void k2_scatter(int n, const int* __restrict__ key,
    const int* __restrict__ rank, const int* __restrict__ sp, const int* __restrict__ offs,
    int* __restrict__ tk, int* __restrict__ tv, int* __restrict__ cnt)
{
    int p = xcd_block((n + BS - 1) / BS) * BS + threadIdx.x;
    if (p >= n) return;
    int k = key[p], r = rank[p];
    int q = offs[k] + r;
    tk[q] = k;
    tv[q] = sp[p];
    if (r == 0) cnt[k] = 0;
}

// [4,5,6,8] final slot = cube start + number of smaller ids in the cube;
// gather of the particle from its old slot curpos[id]; predictor.
__global__ __launch_bounds__(BS) void k2_sortgather(int n, GcParams P, int predict,
    const int* __restrict__ tk, const int* __restrict__ tv, const int* __restrict__ offs,
    int* __restrict__ curpos, const double* __restrict__ pos_old, double* __restrict__ pos_new,
    const float4* __restrict__ vout, const tfloat3* __restrict__ force,
    const tint3* __restrict__ nt_old, tint3* __restrict__ nt_new, int* __restrict__ sp_new,
    int* __restrict__ key_new, float4* __restrict__ posid, tfloat3* __restrict__ coor,
    float4* __restrict__ velrho, tfloat3* __restrict__ velhalf)
{
    int q = xcd_block((n + BS - 1) / BS) * BS + threadIdx.x;
    if (q >= n) return;
    int k = tk[q], id = tv[q];
    int b = offs[k], e = offs[k + 1];
    int r = 0;
    for (int t = b; t < e; ++t) r += tv[t] < id;
    int s = b + r;
    int o = curpos[id];
    double x = pos_old[o], y = pos_old[n + o], z = pos_old[2 * n + o];
    pos_new[s] = x; pos_new[n + s] = y; pos_new[2 * n + s] = z;
    float fx = (float)x, fy = (float)y, fz = (float)z;
    posid[s] = make_float4(fx, fy, fz, __int_as_float(id));
    coor[s] = tfloat3{fx, fy, fz};
    nt_new[s] = nt_old[o];
    sp_new[s] = id;
    key_new[s] = k;
    float4 v = vout[o];
    if (predict && id >= P.nf) {
        tfloat3 f = force[o];
        float h = 0.5f * P.dt * P.inv_m, l = P.lam * P.dt * P.inv_m;
        velhalf[s] = tfloat3{v.x + h * f.x, v.y + h * f.y, v.z + h * f.z};
        v.x += l * f.x; v.y += l * f.y; v.z += l * f.z;
    } else {
        velhalf[s] = tfloat3{v.x, v.y, v.z};
    }
    velrho[s] = v;
    curpos[id] = s;
}

// [7,9] density of moving particles; velocity of Bound2 walls from the
// predicted velocities of their moving neighbours.
__global__ __launch_bounds__(BS) void k2_density(int n, GcParams P, Grid2 g,
    const float4* __restrict__ posid, float4* __restrict__ velrho, const int* __restrict__ offs,
    const int* __restrict__ key)
{
    int p = xcd_block((n + BS - 1) / BS) * BS + threadIdx.x;
    if (p >= n) return;
    float4 pi = posid[p];
    int id = __float_as_int(pi.w);
    if (id < P.nb2) return;
    int cx, cy, cz;
    key_to_cc(key[p], g, cx, cy, cz);
    const float ih = 1.0f / P.h;
    if (id >= P.nf) {
        float d = 0.0f;  // self term added below
        visit2(g, P.per, offs, cx, cy, cz, [&](int jb, int je, float sx, float sy, float sz) {
            for (int j = jb; j < je; ++j) {
                float4 pj = posid[j];
                float rx = pi.x - (pj.x + sx), ry = pi.y - (pj.y + sy), rz = pi.z - (pj.z + sz);
                float r2 = rx * rx + ry * ry + rz * rz;
                if (r2 < P.h2 && __float_as_int(pj.w) >= P.nb2 && j != p) {
                    float q = sqrtf(r2) * ih, om = 1.0f - q;
                    d += (1.0f + 3.0f * q) * om * om * om;
                }
            }
        });
        velrho[p].w = P.m * P.wnorm * (d + 1.0f);
    } else {
        float sw = 0.f, svx = 0.f, svy = 0.f, svz = 0.f;
        visit2(g, P.per, offs, cx, cy, cz, [&](int jb, int je, float sx, float sy, float sz) {
            for (int j = jb; j < je; ++j) {
                float4 pj = posid[j];
                float rx = pi.x - (pj.x + sx), ry = pi.y - (pj.y + sy), rz = pi.z - (pj.z + sz);
                float r2 = rx * rx + ry * ry + rz * rz;
                if (r2 < P.h2 && __float_as_int(pj.w) >= P.nf) {
                    float q = sqrtf(r2) * ih, om = 1.0f - q;
                    float w = (1.0f + 3.0f * q) * om * om * om;
                    float4 vj = velrho[j];
                    sw += w; svx += w * vj.x; svy += w * vj.y; svz += w * vj.z;
                }
            }
        });
        float s = sw > 0.f ? -1.0f / sw : 0.0f;
        float4 v = velrho[p];
        v.x = svx * s; v.y = svy * s; v.z = svz * s;
        velrho[p] = v;
    }
}

// [10-13,17] SDPD pair force with the pressure from the EOS on the fly,
// Bound1 repulsion, body force; corrector for fluid (membrane vertices get
// it after their membrane force).
__global__ __launch_bounds__(BS) void k2_force(int n, GcParams P, Grid2 g, int driveon,
    const float4* __restrict__ posid, const float4* __restrict__ velrho,
    const int* __restrict__ offs, const int* __restrict__ key, const tfloat3* __restrict__ velhalf,
    tfloat3* __restrict__ force, float4* __restrict__ vout, const unsigned long long* __restrict__ ctr)
{
    int p = xcd_block((n + BS - 1) / BS) * BS + threadIdx.x;
    if (p >= n) return;
    float4 pi = posid[p];
    int id = __float_as_int(pi.w);
    if (id < P.nf) { force[p] = tfloat3{0.f, 0.f, 0.f}; return; }
    unsigned long long c = *ctr;
    P.ctr_lo = (uint32_t)c;
    P.ctr_hi = (uint32_t)(c >> 32);
    float4 vi = velrho[p];
    float di = vi.w * P.inv_m;
    float pi_d2 = (id < P.nc ? P.c1sq : P.c2sq) * vi.w / (di * di);
    int cx, cy, cz;
    key_to_cc(key[p], g, cx, cy, cz);
    const float ih = 1.0f / P.h;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    // Pass 1 scans the candidates and keeps the hits (slot << 6 | image
    // code) in this lane's column of an LDS list; pass 2 evaluates them with
    // the wave's lanes in step, so the expensive pair body (EOS, viscous
    // terms, Philox + Box-Muller) runs with nearly all lanes busy instead of
    // the ~15 % that are inside the cut-off at any one candidate.
    __shared__ int hits[FORCE_HMAX * BS];
    int* mine = hits + threadIdx.x;
    int nh = 0;
    auto pair = [&](int code) {
        int j = code >> 6;
        float sx = (code & 3) == 1 ? -g.L[0] : ((code & 3) == 2 ? g.L[0] : 0.0f);
        float sy = ((code >> 2) & 3) == 1 ? -g.L[1] : (((code >> 2) & 3) == 2 ? g.L[1] : 0.0f);
        float sz = ((code >> 4) & 3) == 1 ? -g.L[2] : (((code >> 4) & 3) == 2 ? g.L[2] : 0.0f);
        float4 pj = posid[j];
        float rx = pi.x - (pj.x + sx), ry = pi.y - (pj.y + sy), rz = pi.z - (pj.z + sz);
        float r2 = rx * rx + ry * ry + rz * rz;
        int idj = __float_as_int(pj.w);
        float r = sqrtf(r2);
        float om = 1.0f - r * ih;
        float F = P.fnorm * om * om;
        if (idj < P.nb2) {  // Bound1 wall: repulsion only
            float kk = P.krep * F;
            fx += kk * rx; fy += kk * ry; fz += kk * rz;
            return;
        }
        float4 vj = velrho[j];
        float dj = vj.w * P.inv_m;
        float cC = (pi_d2 + (idj < P.nc ? P.c1sq : P.c2sq) * vj.w / (dj * dj)) * F;
        float ir = 1.0f / r;
        float ex = rx * ir, ey = ry * ir, ez = rz * ir;
        float vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
        float ev = ex * vx + ey * vy + ez * vz;
        float gg = F / (di * dj);
        float ga = gg * P.visc_a, gb = gg * P.visc_b * ev;
        fx += cC * rx - (ga * vx + gb * ex);
        fy += cC * ry - (ga * vy + gb * ey);
        fz += cC * rz - (ga * vz + gb * ez);
        if (P.noise_on) {
            float n0, n1, n2;
            pair_normals(P, (uint32_t)id, (uint32_t)idj, n0, n1, n2);
            float exi = ex * n0 + ey * n1 + ez * n2;
            float amp = sqrtf(P.two_kT_over_dt * gg);
            amp = id < idj ? amp : -amp;
            float pa = amp * P.sq_a, pe = amp * (P.sq_ab - P.sq_a) * exi;
            fx += pa * n0 + pe * ex;
            fy += pa * n1 + pe * ey;
            fz += pa * n2 + pe * ez;
        }
    };
    visit2(g, P.per, offs, cx, cy, cz, [&](int jb, int je, float sx, float sy, float sz) {
        int code = (sx < 0.f ? 1 : (sx > 0.f ? 2 : 0)) | ((sy < 0.f ? 1 : (sy > 0.f ? 2 : 0)) << 2) |
                   ((sz < 0.f ? 1 : (sz > 0.f ? 2 : 0)) << 4);
        for (int j = jb; j < je; ++j) {
            float4 pj = posid[j];
            float rx = pi.x - (pj.x + sx), ry = pi.y - (pj.y + sy), rz = pi.z - (pj.z + sz);
            float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < P.h2 && r2 != 0.f) {
                mine[nh * BS] = (j << 6) | code;
                if (++nh == FORCE_HMAX) {  // list full (dense spot): drain it here
                    for (int k = 0; k < nh; ++k) pair(mine[k * BS]);
                    nh = 0;
                }
            }
        }
    });
    for (int k = 0; k < nh; ++k) pair(mine[k * BS]);
    if (driveon) fx += ext_force_x(P, pi.y);
    force[p] = tfloat3{fx, fy, fz};
    if (id < P.nc) {
        float h = 0.5f * P.dt * P.inv_m;
        tfloat3 vh = velhalf[p];
        vout[p] = make_float4(vh.x + h * fx, vh.y + h * fy, vh.z + h * fz, vi.w);
    }
}

// [16] Morse between membrane vertices of different cells
__global__ __launch_bounds__(BS) void k2_agg(int n, GcParams P, Grid2 g,
    const float4* __restrict__ posid, const int* __restrict__ offs, const int* __restrict__ key,
    const int* __restrict__ celltab, tfloat3* __restrict__ force, tfloat3* __restrict__ aggforce)
{
    int p = xcd_block((n + BS - 1) / BS) * BS + threadIdx.x;
    if (p >= n) return;
    float4 pi = posid[p];
    int id = __float_as_int(pi.w);
    if (id < P.nc) return;
    int cell = celltab[id - P.nc];
    int cx, cy, cz;
    key_to_cc(key[p], g, cx, cy, cz);
    const float tb = 2.0f * P.beta * P.De;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    visit2(g, P.per, offs, cx, cy, cz, [&](int jb, int je, float sx, float sy, float sz) {
        for (int j = jb; j < je; ++j) {
            float4 pj = posid[j];
            int idj = __float_as_int(pj.w);
            if (idj < P.nc) continue;
            float rx = pi.x - (pj.x + sx), ry = pi.y - (pj.y + sy), rz = pi.z - (pj.z + sz);
            float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 >= P.mcut2 || r2 == 0.f || celltab[idj - P.nc] == cell) continue;
            float r = sqrtf(r2);
            float ex = expf(P.beta * (P.r0 - r));
            float fm = tb * (ex * ex - ex) / r;
            fx += fm * rx; fy += fm * ry; fz += fm * rz;
        }
    });
    aggforce[id - P.nc] = tfloat3{fx, fy, fz};
    tfloat3 F = force[p];
    F.x += fx; F.y += fy; F.z += fz;
    force[p] = F;
}

// [18] wall velocity from the corrected velocities; walls' entry of vout
__global__ __launch_bounds__(BS) void k2_wallvel(int n, GcParams P, Grid2 g,
    const float4* __restrict__ posid, float4* __restrict__ vout, const int* __restrict__ offs,
    const int* __restrict__ key)
{
    int p = xcd_block((n + BS - 1) / BS) * BS + threadIdx.x;
    if (p >= n) return;
    float4 pi = posid[p];
    int id = __float_as_int(pi.w);
    if (id >= P.nf) return;
    if (id < P.nb2) { vout[p] = make_float4(0.f, 0.f, 0.f, P.rho0); return; }
    int cx, cy, cz;
    key_to_cc(key[p], g, cx, cy, cz);
    const float ih = 1.0f / P.h;
    float sw = 0.f, svx = 0.f, svy = 0.f, svz = 0.f;
    visit2(g, P.per, offs, cx, cy, cz, [&](int jb, int je, float sx, float sy, float sz) {
        for (int j = jb; j < je; ++j) {
            float4 pj = posid[j];
            float rx = pi.x - (pj.x + sx), ry = pi.y - (pj.y + sy), rz = pi.z - (pj.z + sz);
            float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < P.h2 && __float_as_int(pj.w) >= P.nf) {
                float q = sqrtf(r2) * ih, om = 1.0f - q;
                float w = (1.0f + 3.0f * q) * om * om * om;
                float4 vj = vout[j];
                sw += w; svx += w * vj.x; svy += w * vj.y; svz += w * vj.z;
            }
        }
    });
    float s = sw > 0.f ? -1.0f / sw : 0.0f;
    vout[p] = make_float4(svx * s, svy * s, svz * s, P.rho0);
}

__global__ void k2_bump(unsigned long long* ctr) { *ctr += 1ull; }

// Unwrapped membrane vertex positions in id order (FP64, x[nm] y[nm] z[nm]):
// the membrane force kernel then reads a compact 24 B x Nm array instead of
// gathering coor / nthrobound / curpos from the whole slot-ordered state.
__global__ __launch_bounds__(BS) void k2_mgather(int nm, GcParams P, const int* __restrict__ curpos,
    const tfloat3* __restrict__ coor, const tint3* __restrict__ nthro, const float* __restrict__ bbox,
    double* __restrict__ mpos)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nm) return;
    int s = curpos[P.nc + v];
    tfloat3 x = coor[s];
    tint3 n = nthro[s];
    double L0 = (double)(bbox[1] - bbox[0]), L1 = (double)(bbox[3] - bbox[2]),
           L2 = (double)(bbox[5] - bbox[4]);
    mpos[v] = (double)x.x + (P.per[0] ? (double)n.x * L0 : 0.0);
    mpos[nm + v] = (double)x.y + (P.per[1] ? (double)n.y * L1 : 0.0);
    mpos[2 * nm + v] = (double)x.z + (P.per[2] ? (double)n.z * L2 : 0.0);
}

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

}  // namespace

struct gc_engine {
    int N = 0, nm = 0, ncell = 0, ntri = 0, nedge = 0, K = 0;
    int ncube[3]; float outer[3], invw[3], bbox[6];
    Structconst consts;
    gc_ext_consts ext;
    long step = 0;
    hipStream_t s = nullptr;
    int cur = 0;
    std::vector<void*> allocs;

    float *d_bbox, *d_outer, *d_invw; int* d_ncube;
    double* d_pos64[2]; tint3* d_nt[2]; int* d_sp[2];
    float4 *d_posid, *d_velrho, *d_vout;
    tfloat3 *d_coor, *d_velhalf, *d_force;
    int *d_key0, *d_key, *d_rank, *d_cnt, *d_offs, *d_tk, *d_tv, *d_curpos, *d_scanws;
    unsigned long long* d_ctr;
    // membrane
    tint3* d_tri; int6* d_edge; int* d_celltab; tint2* d_betri;
    float *d_mu0, *d_kc, *d_kd, *d_ka, *d_kv;
    float *d_l0, *d_lmax, *d_kb, *d_mexp, *d_p, *d_kp, *d_th0;
    tfloat3 *d_trinor, *d_tricen; float *d_triarea, *d_reftriarea;
    float *d_refarea_para, *d_refvol_para, *d_area_para, *d_vol_para, *d_area, *d_vol, *d_pararea;
    tfloat3 *d_defforce, *d_aggforce;
    double* d_mpos;
    int *d_es, *d_ee, *d_ts, *d_te;
    hipGraphExec_t gexec = nullptr;
    int g_driveon = -1;
    bool walls_dirty = false;

    template <class T> int alloc(T** p, size_t n)
    {
        int st = dalloc(p, n);
        if (!st) allocs.push_back((void*)*p);
        return st;
    }
};

namespace {

GcParams engine_params(const gc_engine* e)
{
    GcParams P;
    make_params(e->consts, e->ext, &P);
    return P;
}

Grid2 engine_grid(const gc_engine* e)
{
    Grid2 g;
    for (int d = 0; d < 3; ++d) { g.nc[d] = e->ncube[d]; g.L[d] = e->outer[d]; }
    return g;
}

// Neighbour build: keys from the (optionally advanced) FP64 positions,
// counting sort by (cube, id), gather (+ predictor).
int launch_build(gc_engine* e, const GcParams& P, int move, int predict)
{
    const int N = e->N, c = e->cur, nx = 1 - c;
    hipStream_t s = e->s;
    const int gb = blocks_for(N, BS);
    k2_integrate<<<xcd_grid(gb), BS, 0, s>>>(N, P, move, e->d_pos64[c], e->d_vout, e->d_force, e->d_sp[c],
        e->d_nt[c], e->d_bbox, e->d_invw, e->d_ncube, e->d_key0, e->d_rank, e->d_cnt);
    GC_LAUNCH_CHECK("engine integrate");
    int st = exclusive_scan(e->d_cnt, e->d_offs, e->K, s, e->d_scanws);
    if (st) return st;
    k2_scatter<<<xcd_grid(gb), BS, 0, s>>>(N, e->d_key0, e->d_rank, e->d_sp[c], e->d_offs, e->d_tk, e->d_tv,
                                 e->d_cnt);
    GC_LAUNCH_CHECK("engine scatter");
    k2_sortgather<<<xcd_grid(gb), BS, 0, s>>>(N, P, predict, e->d_tk, e->d_tv, e->d_offs, e->d_curpos,
        e->d_pos64[c], e->d_pos64[nx], e->d_vout, e->d_force, e->d_nt[c], e->d_nt[nx], e->d_sp[nx],
        e->d_key, e->d_posid, e->d_coor, e->d_velrho, e->d_velhalf);
    GC_LAUNCH_CHECK("engine sort/gather");
    e->cur = nx;
    return GC_OK;
}

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

int launch_step(gc_engine* e, int driveon, Prof* pr)
{
    GcParams P = engine_params(e);
    Grid2 g = engine_grid(e);
    const int N = e->N;
    hipStream_t s = e->s;
    const int gb = blocks_for(N, BS);
    mark(pr, s, "start");
    int st = launch_build(e, P, 1, 1);
    if (st) return st;
    mark(pr, s, "neighbour build+predict");
    const int c = e->cur;
    k2_density<<<xcd_grid(gb), BS, 0, s>>>(N, P, g, e->d_posid, e->d_velrho, e->d_offs, e->d_key);
    GC_LAUNCH_CHECK("engine density");
    mark(pr, s, "density+wall velocity");
    k2_force<<<xcd_grid(gb), BS, 0, s>>>(N, P, g, driveon, e->d_posid, e->d_velrho, e->d_offs, e->d_key,
                               e->d_velhalf, e->d_force, e->d_vout, e->d_ctr);
    GC_LAUNCH_CHECK("engine force");
    k2_bump<<<1, 1, 0, s>>>(e->d_ctr);
    mark(pr, s, "sdpd force");
    if (e->ncell > 0) {
        k_geo_cell<<<e->ncell, BS, 0, s>>>(P, e->step, e->d_bbox, e->d_curpos, e->d_tri, e->d_betri,
            e->d_nt[c], e->d_coor, e->d_trinor, e->d_triarea, e->d_reftriarea, e->d_tricen,
            e->d_refarea_para, e->d_refvol_para, e->d_area_para, e->d_vol_para, e->d_area, e->d_vol);
        GC_LAUNCH_CHECK("engine geo");
        mark(pr, s, "membrane geometry");
        if (P.De != 0.0f) {
            k2_agg<<<xcd_grid(gb), BS, 0, s>>>(N, P, g, e->d_posid, e->d_offs, e->d_key, e->d_celltab,
                                     e->d_force, e->d_aggforce);
            GC_LAUNCH_CHECK("engine morse");
        }
        mark(pr, s, "morse");
        DefArgs a;
        a.edge = e->d_edge; a.tri = e->d_tri; a.coor = e->d_coor; a.nthro = e->d_nt[c];
        a.curpos = e->d_curpos; a.lmax = e->d_lmax; a.mexp = e->d_mexp; a.p = e->d_p;
        a.kp = e->d_kp; a.th0 = e->d_th0; a.kb = e->d_kb; a.reftriarea = e->d_reftriarea;
        a.kd = e->d_kd; a.ka = e->d_ka; a.Ac = e->d_area_para; a.A0c = e->d_refarea_para;
        a.kv = e->d_kv; a.Vc = e->d_vol_para; a.V0c = e->d_refvol_para;
        a.es = e->d_es; a.ee = e->d_ee; a.ts = e->d_ts; a.te = e->d_te;
        a.defforce = e->d_defforce; a.force = e->d_force;
        a.velhalf = e->d_velhalf; a.velrho = e->d_velrho; a.vout = e->d_vout;
        a.mpos = e->d_mpos; a.nm = e->nm;
        k2_mgather<<<blocks_for(e->nm, BS), BS, 0, s>>>(e->nm, P, e->d_curpos, e->d_coor, e->d_nt[c],
                                                        e->d_bbox, e->d_mpos);
        GC_LAUNCH_CHECK("engine membrane gather");
        // 64-thread groups: a few hundred RBCs give only ~Nm/256 groups of 256
        k_def_force<<<xcd_grid(blocks_for(e->nm, 64)), 64, 0, s>>>(P, e->nm, e->d_bbox, a);
        GC_LAUNCH_CHECK("engine membrane force");
        mark(pr, s, "membrane force+corrector");
    }
    // [18] CalBound2Vel after the corrector only feeds the output: the next
    // step overwrites the walls' velocity (k2_density) before anything reads
    // it.  It is computed when the state leaves the engine (finish_walls).
    e->walls_dirty = P.nf > 0;
    e->step += 1;
    return GC_OK;
}

// Wall velocities from the corrected flow (UpdatePar's final CalBound2Vel,
// Particle.cu:1327) for the exported state.
int finish_walls(gc_engine* e)
{
    if (!e->walls_dirty) return GC_OK;
    GcParams P = engine_params(e);
    Grid2 g = engine_grid(e);
    k2_wallvel<<<xcd_grid(blocks_for(e->N, BS)), BS, 0, e->s>>>(e->N, P, g, e->d_posid, e->d_vout, e->d_offs,
                                                        e->d_key);
    GC_LAUNCH_CHECK("engine wall velocity");
    e->walls_dirty = false;
    return GC_OK;
}

void destroy(gc_engine* e)
{
    if (!e) return;
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->s) (void)hipStreamSynchronize(e->s);
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

}  // namespace

extern "C" gc_engine* gc_EngineCreate(const gc_case_desc* d, int* status)
{
    int dummy;
    if (!status) status = &dummy;
    *status = GC_OK;
    if (!d || !d->coor || d->nBound1Par < 0 || d->nBound2Par < 0 || d->nFluidPar < 0 ||
        d->nCell < 0 || d->nCellPar < 0) {
        *status = set_error(GC_ERR_INVALID, "EngineCreate: bad descriptor");
        return nullptr;
    }
    if (d->nCell > 0 && (!d->tri || !d->edge || !d->celltab || !d->beginendtri ||
                         !d->shearmod_para || !d->bendingmod_para || !d->hlarea_para ||
                         !d->hgarea_para || !d->hvol_para || d->nAllTri <= 0 ||
                         d->nAllEdge <= 0 || d->nCellPar <= 0)) {
        *status = set_error(GC_ERR_INVALID, "EngineCreate: membrane arrays missing");
        return nullptr;
    }
    gc_engine* e = new gc_engine();
    int st = GC_OK;
    auto fail = [&](int code) { *status = code; destroy(e); return (gc_engine*)nullptr; };

    e->consts = d->consts;
    e->ext = d->ext;
    e->consts.nBound2Start = (unsigned)d->nBound1Par;
    e->consts.nFluidStart = (unsigned)(d->nBound1Par + d->nBound2Par);
    e->consts.nCellStart = e->consts.nFluidStart + (unsigned)d->nFluidPar;
    e->consts.nCellPar = (unsigned)(d->nCell > 0 ? d->nCellPar : 0);
    e->consts.nTotalPar = e->consts.nCellStart + e->consts.nCellPar;
    if (!(e->consts.CutRadius > 0) || !(e->consts.Mass > 0) || !(e->consts.TimeStep > 0) ||
        !(e->consts.DenInit > 0))
        return fail(set_error(GC_ERR_INVALID, "EngineCreate: CutRadius, Mass, TimeStep, DenInit must be > 0"));
    if (5.0 * e->consts.ShearVis / 3.0 - e->consts.BulkVis < 0)
        return fail(set_error(GC_ERR_INVALID, "EngineCreate: need 5*ShearVis/3 >= BulkVis"));
    if (e->consts.surEnergy != 0 && e->consts.MaxCut > e->consts.CutRadius)
        return fail(set_error(GC_ERR_GRID, "EngineCreate: MaxCut must not exceed CutRadius"));
    if (!(e->ext.wlc_ratio > 1.0f))
        return fail(set_error(GC_ERR_INVALID, "EngineCreate: wlc_ratio must be > 1"));

    e->N = (int)e->consts.nTotalPar;
    e->nm = (int)e->consts.nCellPar;
    e->ncell = d->nCell;
    e->ntri = d->nCell > 0 ? d->nAllTri : 0;
    e->nedge = d->nCell > 0 ? d->nAllEdge : 0;
    memcpy(e->bbox, d->bbox, sizeof(e->bbox));
    if ((st = gc_ConfigDomain(e->bbox, e->consts.CutRadius, e->ext.periodic, e->ncube, e->outer,
                              e->invw, &e->K)))
        return fail(st);
    if ((st = hip_check(hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking), "engine stream")))
        return fail(st);

    const size_t N = (size_t)e->N, K = (size_t)e->K;
    if (e->N >= (1 << 25))  // the force kernel packs slot << 6 into an int
        return fail(set_error(GC_ERR_INVALID, "EngineCreate: more than 2^25 particles"));
    if ((st = e->alloc(&e->d_bbox, 6)) || (st = e->alloc(&e->d_outer, 3)) ||
        (st = e->alloc(&e->d_invw, 3)) || (st = e->alloc(&e->d_ncube, 3)))
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
        (st = e->alloc(&e->d_tv, N)) || (st = e->alloc(&e->d_curpos, N)) ||
        (st = e->alloc(&e->d_scanws, scan_ws_ints(e->K))) || (st = e->alloc(&e->d_ctr, 1)))
        return fail(st);

    hipStream_t s = e->s;
    std::vector<float4> vout_id(N);
    std::vector<tfloat3> vh_id(N), f_id(N, tfloat3{0.f, 0.f, 0.f});
    {
        // Particle::SetParticle (Particle.cu:485-546): rho = DenInit, F = 0
        std::vector<double> pos(3 * N);
        std::vector<tint3> nt(N, tint3{0, 0, 0});
        std::vector<int> idn(N), zero(K + 1, 0);
        for (size_t i = 0; i < N; ++i) {
            pos[i] = d->coor[i].x; pos[N + i] = d->coor[i].y; pos[2 * N + i] = d->coor[i].z;
            tfloat3 v = d->vel ? d->vel[i] : tfloat3{0.f, 0.f, 0.f};
            if ((int)i < (int)e->consts.nBound2Start) v = tfloat3{0.f, 0.f, 0.f};
            vout_id[i] = make_float4(v.x, v.y, v.z, e->consts.DenInit);
            vh_id[i] = v;
            idn[i] = (int)i;
        }
        unsigned long long ctr = e->ext.noise_counter;
        if ((st = upload(e->d_bbox, e->bbox, 6, s)) || (st = upload(e->d_outer, e->outer, 3, s)) ||
            (st = upload(e->d_invw, e->invw, 3, s)) || (st = upload(e->d_ncube, e->ncube, 3, s)) ||
            (st = upload(e->d_pos64[0], pos.data(), 3 * N, s)) ||
            (st = upload(e->d_vout, vout_id.data(), N, s)) || (st = upload(e->d_force, f_id.data(), N, s)) ||
            (st = upload(e->d_nt[0], nt.data(), N, s)) || (st = upload(e->d_sp[0], idn.data(), N, s)) ||
            (st = upload(e->d_curpos, idn.data(), N, s)) || (st = upload(e->d_cnt, zero.data(), K + 1, s)) ||
            (st = upload(e->d_ctr, &ctr, 1, s)))
            return fail(st);
    }
    GcParams P = engine_params(e);
    // first neighbour build (Particle::ConCubeList at init, Particle.cu:570)
    if ((st = launch_build(e, P, 0, 0))) return fail(st);
    if ((st = reorder_carried(e, vout_id, f_id, vh_id))) return fail(st);

    if (e->ncell > 0) {
        const size_t T = (size_t)e->ntri, E = (size_t)e->nedge, C = (size_t)e->ncell, M = (size_t)e->nm;
        if ((st = e->alloc(&e->d_tri, T)) || (st = e->alloc(&e->d_edge, E)) ||
            (st = e->alloc(&e->d_celltab, M)) || (st = e->alloc(&e->d_betri, C)) ||
            (st = e->alloc(&e->d_mu0, E)) || (st = e->alloc(&e->d_kc, E)) ||
            (st = e->alloc(&e->d_kd, T)) || (st = e->alloc(&e->d_ka, T)) || (st = e->alloc(&e->d_kv, T)) ||
            (st = e->alloc(&e->d_l0, E)) || (st = e->alloc(&e->d_lmax, E)) || (st = e->alloc(&e->d_kb, E)) ||
            (st = e->alloc(&e->d_mexp, E)) || (st = e->alloc(&e->d_p, E)) || (st = e->alloc(&e->d_kp, E)) ||
            (st = e->alloc(&e->d_th0, E)) || (st = e->alloc(&e->d_trinor, T)) ||
            (st = e->alloc(&e->d_tricen, T)) || (st = e->alloc(&e->d_triarea, T)) ||
            (st = e->alloc(&e->d_reftriarea, T)) || (st = e->alloc(&e->d_refarea_para, T)) ||
            (st = e->alloc(&e->d_refvol_para, T)) || (st = e->alloc(&e->d_area_para, T)) ||
            (st = e->alloc(&e->d_vol_para, T)) || (st = e->alloc(&e->d_area, C)) ||
            (st = e->alloc(&e->d_vol, C)) || (st = e->alloc(&e->d_pararea, M)) ||
            (st = e->alloc(&e->d_defforce, M)) || (st = e->alloc(&e->d_aggforce, M)) ||
            (st = e->alloc(&e->d_mpos, 3 * M)))
            return fail(st);
        std::vector<int> es, ee, ts, te;
        if ((st = build_incidence(e->nm, e->nedge, d->edge, e->ntri, d->tri, es, ee, ts, te)))
            return fail(st);
        for (int c = 0; c < e->ncell; ++c)
            if (d->beginendtri[c].x < 0 || d->beginendtri[c].y >= e->ntri ||
                d->beginendtri[c].x > d->beginendtri[c].y)
                return fail(set_error(GC_ERR_TOPOLOGY, "EngineCreate: bad BeginEndTri"));
        for (int v = 0; v < e->nm; ++v)
            if (d->celltab[v] < 0 || d->celltab[v] >= e->ncell)
                return fail(set_error(GC_ERR_TOPOLOGY, "EngineCreate: CellTab out of range"));
        if ((st = e->alloc(&e->d_es, es.size())) || (st = e->alloc(&e->d_ee, ee.size())) ||
            (st = e->alloc(&e->d_ts, ts.size())) || (st = e->alloc(&e->d_te, te.size())))
            return fail(st);
        std::vector<tfloat3> zero(M, tfloat3{0.f, 0.f, 0.f});
        if ((st = upload(e->d_tri, d->tri, T, s)) || (st = upload(e->d_edge, d->edge, E, s)) ||
            (st = upload(e->d_celltab, d->celltab, M, s)) || (st = upload(e->d_betri, d->beginendtri, C, s)) ||
            (st = upload(e->d_mu0, d->shearmod_para, E, s)) || (st = upload(e->d_kc, d->bendingmod_para, E, s)) ||
            (st = upload(e->d_kd, d->hlarea_para, T, s)) || (st = upload(e->d_ka, d->hgarea_para, T, s)) ||
            (st = upload(e->d_kv, d->hvol_para, T, s)) || (st = upload(e->d_es, es.data(), es.size(), s)) ||
            (st = upload(e->d_ee, ee.data(), ee.size(), s)) || (st = upload(e->d_ts, ts.data(), ts.size(), s)) ||
            (st = upload(e->d_te, te.data(), te.size(), s)) || (st = upload(e->d_defforce, zero.data(), M, s)) ||
            (st = upload(e->d_aggforce, zero.data(), M, s)))
            return fail(st);
        const int c = e->cur;
        // reference geometry and mechanics (Particle::SetPar, Particle.cu:476-479)
        k_geo_cell<<<e->ncell, BS, 0, s>>>(P, -1L, e->d_bbox, e->d_curpos, e->d_tri, e->d_betri,
            e->d_nt[c], e->d_coor, e->d_trinor, e->d_triarea, e->d_reftriarea, e->d_tricen,
            e->d_refarea_para, e->d_refvol_para, e->d_area_para, e->d_vol_para, e->d_area, e->d_vol);
        k_par_area<<<blocks_for(e->nm, BS), BS, 0, s>>>(e->nm, e->d_ts, e->d_te, e->d_triarea, e->d_pararea);
        k_mech_para<<<blocks_for(e->nedge, BS), BS, 0, s>>>(P, e->nedge, e->d_edge, e->d_curpos,
            e->d_coor, e->d_l0, e->d_lmax, e->d_kc, e->d_mu0, e->d_kb, e->d_mexp, e->d_p, e->d_kp,
            e->d_th0);
        if ((st = hip_check(hipGetLastError(), "engine membrane init"))) return fail(st);
    }
    if ((st = hip_check(hipStreamSynchronize(s), "engine init sync"))) return fail(st);
    e->step = 0;
    return e;
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
    if (nsteps & 1) return launch_step(e, driveon, nullptr);
    return GC_OK;
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
    info->nTotalPar = e->N; info->nCellPar = e->nm; info->nCell = e->ncell; info->sizecube = e->K;
    for (int d = 0; d < 3; ++d) {
        info->ncube[d] = e->ncube[d]; info->outerregion[d] = e->outer[d]; info->invwid[d] = e->invw[d];
    }
    return GC_OK;
}

extern "C" int gc_EngineGetState(gc_engine* e, gc_host_state* o)
{
    if (!e || !o) return set_error(GC_ERR_INVALID, "EngineGetState: null");
    const int N = e->N, c = e->cur;
    int st = finish_walls(e);
    if (st) return st;
    st = hip_check(hipStreamSynchronize(e->s), "GetState sync");
    if (st) return st;
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
    const float c1sq = e->consts.SoundSpeed1 * e->consts.SoundSpeed1;
    const float c2sq = e->consts.SoundSpeed2 * e->consts.SoundSpeed2;
    for (int slot = 0; slot < N; ++slot) {
        int id = sp[slot];
        if (o->pos) for (int d = 0; d < 3; ++d) o->pos[3 * (size_t)id + d] = pos[(size_t)d * N + slot];
        if (o->coor) o->coor[id] = co[slot];
        if (o->velrhop) o->velrhop[id] = tfloat4{vo[slot].x, vo[slot].y, vo[slot].z, vo[slot].w};
        if (o->velhalf) o->velhalf[id] = vh[slot];
        if (o->force) o->force[id] = f[slot];
        if (o->press) o->press[id] = (id < (int)e->consts.nCellStart ? c1sq : c2sq) * vo[slot].w;
        if (o->nthrobound) o->nthrobound[id] = nt[slot];
    }
    if (e->ncell > 0) {
        if (o->area && (st = dl(o->area, e->d_area, sizeof(float) * e->ncell))) return st;
        if (o->vol && (st = dl(o->vol, e->d_vol, sizeof(float) * e->ncell))) return st;
        if (o->defforce && (st = dl(o->defforce, e->d_defforce, sizeof(tfloat3) * e->nm))) return st;
        if (o->aggforce && (st = dl(o->aggforce, e->d_aggforce, sizeof(tfloat3) * e->nm))) return st;
    }
    return GC_OK;
}

extern "C" int gc_EngineSetState(gc_engine* e, const gc_host_state* in, long step,
                                 uint64_t noise_counter)
{
    if (!e || !in) return set_error(GC_ERR_INVALID, "EngineSetState: null");
    const int N = e->N;
    int st = hip_check(hipStreamSynchronize(e->s), "SetState sync");
    if (st) return st;
    if (e->gexec) { (void)hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
    std::vector<double> pos(3 * (size_t)N);
    std::vector<tfloat4> vr(N);
    std::vector<tfloat3> vh(N), f(N);
    std::vector<tint3> nt(N);
    gc_host_state cur;
    memset(&cur, 0, sizeof(cur));
    cur.pos = pos.data(); cur.velrhop = vr.data(); cur.velhalf = vh.data(); cur.force = f.data();
    cur.nthrobound = nt.data();
    if ((st = gc_EngineGetState(e, &cur))) return st;
    if (in->pos) memcpy(pos.data(), in->pos, sizeof(double) * 3 * N);
    std::vector<double> pos_soa(3 * (size_t)N);
    for (int i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) pos_soa[(size_t)d * N + i] = pos[3 * (size_t)i + d];
    if (in->velrhop) memcpy(vr.data(), in->velrhop, sizeof(tfloat4) * N);
    if (in->velhalf) memcpy(vh.data(), in->velhalf, sizeof(tfloat3) * N);
    if (in->force) memcpy(f.data(), in->force, sizeof(tfloat3) * N);
    if (in->nthrobound) memcpy(nt.data(), in->nthrobound, sizeof(tint3) * N);
    std::vector<int> idn(N);
    std::vector<float4> vo(N);
    for (int i = 0; i < N; ++i) { idn[i] = i; vo[i] = make_float4(vr[i].x, vr[i].y, vr[i].z, vr[i].w); }
    e->cur = 0;
    hipStream_t s = e->s;
    unsigned long long ctr = noise_counter;
    if ((st = upload(e->d_pos64[0], pos_soa.data(), 3 * (size_t)N, s)) || (st = upload(e->d_vout, vo.data(), N, s)) ||
        (st = upload(e->d_force, f.data(), N, s)) || (st = upload(e->d_nt[0], nt.data(), N, s)) ||
        (st = upload(e->d_sp[0], idn.data(), N, s)) || (st = upload(e->d_curpos, idn.data(), N, s)) ||
        (st = upload(e->d_ctr, &ctr, 1, s)))
        return st;
    GcParams P = engine_params(e);
    if ((st = launch_build(e, P, 0, 0))) return st;
    if ((st = reorder_carried(e, vo, f, vh))) return st;
    e->step = step;
    return GC_OK;
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
    std::vector<int> key(N), offs(K + 1);
    if ((st = dl(o->sortpart, e->d_sp[c], sizeof(int) * N)) ||
        (st = dl(o->curpos, e->d_curpos, sizeof(int) * N)) ||
        (st = dl(key.data(), e->d_key, sizeof(int) * N)) ||
        (st = dl(offs.data(), e->d_offs, sizeof(int) * (K + 1))))
        return st;
    if (o->cubepart) memcpy(o->cubepart, key.data(), sizeof(int) * N);
    if (o->cc)
        for (int s = 0; s < N; ++s) {
            int k = key[s];
            o->cc[s] = tint3{k % e->ncube[0], (k / e->ncube[0]) % e->ncube[1], k / (e->ncube[0] * e->ncube[1])};
        }
    if (o->beginendcube)
        for (int k = 0; k < K; ++k) o->beginendcube[k] = tint2{offs[k], offs[k + 1]};
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
