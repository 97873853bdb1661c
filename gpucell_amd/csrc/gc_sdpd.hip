// gc_sdpd.hip -- fluid-path entry points replacing cupar::UpdateFluidCoor,
// AppFluidPeriodicBC, CalFluidDen, PredictFluidVel, CalBound2Vel, EOS,
// CalFluidSDPDForce, CalFluidRepForce, SetFluidExtForce and CorrectFliudVel
// (src/ParticleGPU.cuh:11-21).  Arrays are slot-indexed (sorted order) in
// the reference's TypesDef layouts; the particle kind comes from its id
// (sortpart[slot]) and the id ranges of the parameter block.
//
// Physics (Lucy kernel support h = CutRadius, number density d = rho/m):
//   rho_i = m sum_j W(r_ij)                               (self included)
//   P_i   = c_i^2 rho_i
//   f_ij  = (P_i/d_i^2 + P_j/d_j^2) F r_ij
//           - F/(d_i d_j) [a v_ij + b e (e.v_ij)]
//           + s_ij sqrt(2 kT F/(d_i d_j)/dt) [sqrt(a)(xi - e(e.xi)) + sqrt(a+b) e(e.xi)]
//   a = 5 eta/3 - zeta, b = 5 (zeta + eta/3); F(r) = -W'(r)/r;
//   xi: three Philox-4x32-10 normals of the pair, s_ij = +1 for the lower id.
#include "gc_kernels.h"

using namespace gc;

namespace gc {

constexpr int BS = GC_BS;

__device__ __forceinline__ GcGridV load_grid(const int* ncube, const float* outer)
{
    GcGridV g;
    g.nc[0] = ncube[0]; g.nc[1] = ncube[1]; g.nc[2] = ncube[2];
    g.L[0] = outer[0]; g.L[1] = outer[1]; g.L[2] = outer[2];
    return g;
}

__global__ __launch_bounds__(BS) void k_update_coor(int n, GcParams P, const int* __restrict__ sortpart,
                                                    const tfloat3* __restrict__ force,
                                                    tfloat3* __restrict__ coor,
                                                    const tfloat4* __restrict__ velrhop)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || sortpart[p] < P.nf) return;
    float a = 0.5f * P.dt * P.dt * P.inv_m;
    tfloat3 x = coor[p], f = force[p];
    tfloat4 v = velrhop[p];
    x.x += P.dt * v.x + a * f.x;
    x.y += P.dt * v.y + a * f.y;
    x.z += P.dt * v.z + a * f.z;
    coor[p] = x;
}

__global__ __launch_bounds__(BS) void k_periodic(int n, GcParams P, const int* __restrict__ sortpart,
                                                 tfloat3* __restrict__ coor,
                                                 tint3* __restrict__ nthro,
                                                 const float* __restrict__ bbox)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || sortpart[p] < P.nf) return;
    float x[3] = {coor[p].x, coor[p].y, coor[p].z};
    int img[3] = {nthro[p].x, nthro[p].y, nthro[p].z};
    bool moved = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!P.per[d]) continue;
        float lo = bbox[2 * d], hi = bbox[2 * d + 1], L = hi - lo;
        if (x[d] < lo) { x[d] += L; img[d] -= 1; moved = true; }
        else if (x[d] >= hi) { x[d] -= L; img[d] += 1; moved = true; }
    }
    if (moved) {
        coor[p] = tfloat3{x[0], x[1], x[2]};
        nthro[p] = tint3{img[0], img[1], img[2]};
    }
}

__global__ __launch_bounds__(BS) void k_density(int n, GcParams P, const tfloat3* __restrict__ coor,
                                                tfloat4* __restrict__ velrhop,
                                                const tint3* __restrict__ cc,
                                                const int* __restrict__ ncube,
                                                const float* __restrict__ outer,
                                                const int* __restrict__ sortpart,
                                                const tint2* __restrict__ be)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int id = sortpart[p];
    if (id < P.nf) return;
    GcGridV g = load_grid(ncube, outer);
    tfloat3 xi = coor[p];
    tint3 c = cc[p];
    float d = P.wnorm;  // self term W(0)
    visit_ranges(g, P.per, be, c.x, c.y, c.z, [&](int jb, int je, float sx, float sy, float sz) {
        for (int j = jb; j < je; ++j) {
            if (j == p || sortpart[j] < P.nb2) continue;
            tfloat3 xj = coor[j];
            float rx = pdiff(xi.x, xj.x, sx), ry = pdiff(xi.y, xj.y, sy), rz = pdiff(xi.z, xj.z, sz);
            float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < P.h2) d += lucy_w(sqrtf(r2), P);
        }
    });
    velrhop[p].w = P.m * d;
}

__global__ __launch_bounds__(BS) void k_bound2_vel(int n, GcParams P, const tfloat3* __restrict__ coor,
                                                   tfloat4* __restrict__ velrhop,
                                                   const tint3* __restrict__ cc,
                                                   const int* __restrict__ ncube,
                                                   const float* __restrict__ outer,
                                                   const int* __restrict__ sortpart,
                                                   const tint2* __restrict__ be)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int id = sortpart[p];
    if (id < P.nb2 || id >= P.nf) return;
    GcGridV g = load_grid(ncube, outer);
    tfloat3 xi = coor[p];
    tint3 c = cc[p];
    float sw = 0.f, svx = 0.f, svy = 0.f, svz = 0.f;
    visit_ranges(g, P.per, be, c.x, c.y, c.z, [&](int jb, int je, float sx, float sy, float sz) {
        for (int j = jb; j < je; ++j) {
            if (sortpart[j] < P.nf) continue;
            tfloat3 xj = coor[j];
            float rx = pdiff(xi.x, xj.x, sx), ry = pdiff(xi.y, xj.y, sy), rz = pdiff(xi.z, xj.z, sz);
            float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < P.h2) {
                float w = lucy_w(sqrtf(r2), P);
                tfloat4 vj = velrhop[j];
                sw += w; svx += w * vj.x; svy += w * vj.y; svz += w * vj.z;
            }
        }
    });
    const float3 u = wall_velocity(P, xi.y, sw, svx, svy, svz);
    tfloat4 v = velrhop[p];
    v.x = u.x; v.y = u.y; v.z = u.z;
    velrhop[p] = v;
}

__global__ __launch_bounds__(BS) void k_eos(int n, GcParams P, const int* __restrict__ sortpart,
                                            const tfloat4* __restrict__ velrhop,
                                            float* __restrict__ press)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int id = sortpart[p];
    press[p] = (id < P.nc ? P.c1sq : P.c2sq) * velrhop[p].w;
}

__global__ __launch_bounds__(BS) void k_sdpd_force(int n, GcParams P, const tfloat3* __restrict__ coor,
                                                   const tfloat4* __restrict__ velrhop,
                                                   const tint3* __restrict__ cc,
                                                   const int* __restrict__ ncube,
                                                   const int* __restrict__ sortpart,
                                                   const float* __restrict__ outer,
                                                   const tint2* __restrict__ be,
                                                   tfloat3* __restrict__ force,
                                                   const float* __restrict__ press,
                                                   const unsigned long long* __restrict__ ctr)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int id = sortpart[p];
    if (id < P.nf) { force[p] = tfloat3{0.f, 0.f, 0.f}; return; }
    if (ctr) {  // counter kept in device memory (graph replays advance it)
        unsigned long long c = *ctr;
        P.ctr_lo = (uint32_t)c;
        P.ctr_hi = (uint32_t)(c >> 32);
    }
    GcGridV g = load_grid(ncube, outer);
    tfloat3 xi = coor[p];
    tfloat4 vi = velrhop[p];
    tint3 c = cc[p];
    float di = vi.w * P.inv_m;
    float pi_d2 = press[p] / (di * di);
    float fx = 0.f, fy = 0.f, fz = 0.f;
    visit_ranges(g, P.per, be, c.x, c.y, c.z, [&](int jb, int je, float sx, float sy, float sz) {
        for (int j = jb; j < je; ++j) {
            if (j == p) continue;
            int idj = sortpart[j];
            if (idj < P.nb2) continue;
            tfloat3 xj = coor[j];
            float rx = pdiff(xi.x, xj.x, sx), ry = pdiff(xi.y, xj.y, sy), rz = pdiff(xi.z, xj.z, sz);
            float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 >= P.h2 || r2 == 0.f) continue;
            float r = sqrtf(r2);
            float F = lucy_f(r, P);
            tfloat4 vj = velrhop[j];
            float dj = vj.w * P.inv_m;
            float cC = (pi_d2 + press[j] / (dj * dj)) * F;
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
        }
    });
    force[p] = tfloat3{fx, fy, fz};
}

__global__ __launch_bounds__(BS) void k_rep_force(int n, GcParams P, const tfloat3* __restrict__ coor,
                                                  const tint3* __restrict__ cc,
                                                  const int* __restrict__ ncube,
                                                  const int* __restrict__ sortpart,
                                                  const float* __restrict__ outer,
                                                  const tint2* __restrict__ be,
                                                  tfloat3* __restrict__ force)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (sortpart[p] < P.nf) return;
    GcGridV g = load_grid(ncube, outer);
    tfloat3 xi = coor[p];
    tint3 c = cc[p];
    float fx = 0.f, fy = 0.f, fz = 0.f;
    visit_ranges(g, P.per, be, c.x, c.y, c.z, [&](int jb, int je, float sx, float sy, float sz) {
        for (int j = jb; j < je; ++j) {
            if (sortpart[j] >= P.nb2) continue;
            tfloat3 xj = coor[j];
            float rx = pdiff(xi.x, xj.x, sx), ry = pdiff(xi.y, xj.y, sy), rz = pdiff(xi.z, xj.z, sz);
            float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 >= P.h2 || r2 == 0.f) continue;
            float k = P.krep * lucy_f(sqrtf(r2), P);
            fx += k * rx; fy += k * ry; fz += k * rz;
        }
    });
    tfloat3 f = force[p];
    f.x += fx; f.y += fy; f.z += fz;
    force[p] = f;
}

__global__ __launch_bounds__(BS) void k_ext_force(int n, GcParams P, const int* __restrict__ sortpart,
                                                  const tfloat3* __restrict__ coor,
                                                  tfloat3* __restrict__ force)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || sortpart[p] < P.nf) return;
    float f = P.fext;
    if (P.force_mode == GC_FORCE_REVERSE_POISEUILLE && coor[p].y >= P.force_split) f = -f;
    force[p].x += f;
}

__global__ __launch_bounds__(BS) void k_predict(int n, GcParams P, const int* __restrict__ sortpart,
                                                tfloat4* __restrict__ velrhop,
                                                tfloat3* __restrict__ velhalf,
                                                const tfloat3* __restrict__ force)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || sortpart[p] < P.nf) return;
    tfloat4 v = velrhop[p];
    tfloat3 f = force[p];
    float h = 0.5f * P.dt * P.inv_m, l = P.lam * P.dt * P.inv_m;
    velhalf[p] = tfloat3{v.x + h * f.x, v.y + h * f.y, v.z + h * f.z};
    v.x += l * f.x; v.y += l * f.y; v.z += l * f.z;
    velrhop[p] = v;
}

__global__ __launch_bounds__(BS) void k_correct(int n, GcParams P, const int* __restrict__ sortpart,
                                                const tfloat3* __restrict__ force,
                                                const tfloat3* __restrict__ velhalf,
                                                tfloat4* __restrict__ velrhop)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || sortpart[p] < P.nf) return;
    tfloat3 f = force[p], vh = velhalf[p];
    float h = 0.5f * P.dt * P.inv_m;
    tfloat4 v = velrhop[p];
    v.x = vh.x + h * f.x; v.y = vh.y + h * f.y; v.z = vh.z + h * f.z;
    velrhop[p] = v;
}

}  // namespace gc

#define GC_PARAMS_OR_RETURN(P)          \
    GcParams P;                         \
    do { int _s = params(&P); if (_s) return _s; } while (0)

#define GC_REQUIRE(cond, msg) \
    do { if (!(cond)) return set_error(GC_ERR_INVALID, msg); } while (0)

extern "C" int gc_UpdateFluidCoor(int n, int* d_sortpart, tfloat3* d_totalforce,
                                  tfloat3* d_totalcoor, tfloat4* d_velrhop)
{
    GC_PARAMS_OR_RETURN(P);
    GC_REQUIRE(n >= 0, "UpdateFluidCoor: n < 0");
    if (n == 0) return GC_OK;
    GC_REQUIRE(d_sortpart && d_totalforce && d_totalcoor && d_velrhop, "UpdateFluidCoor: null pointer");
    k_update_coor<<<blocks_for(n, BS), BS, 0, stream()>>>(n, P, d_sortpart, d_totalforce,
                                                         d_totalcoor, d_velrhop);
    GC_LAUNCH_CHECK("UpdateFluidCoor");
    return GC_OK;
}

extern "C" int gc_AppFluidPeriodicBC(int n, int* d_sortpart, tfloat3* d_totalcoor,
                                     tint3* d_nthrobound, float* d_bboxsdpd)
{
    GC_PARAMS_OR_RETURN(P);
    GC_REQUIRE(n >= 0, "AppFluidPeriodicBC: n < 0");
    if (n == 0) return GC_OK;
    GC_REQUIRE(d_sortpart && d_totalcoor && d_nthrobound && d_bboxsdpd, "AppFluidPeriodicBC: null pointer");
    k_periodic<<<blocks_for(n, BS), BS, 0, stream()>>>(n, P, d_sortpart, d_totalcoor,
                                                      d_nthrobound, d_bboxsdpd);
    GC_LAUNCH_CHECK("AppFluidPeriodicBC");
    return GC_OK;
}

extern "C" int gc_CalFluidDen(int n, tfloat3* d_totalcoor, tfloat4* d_velrhop, tint3* d_cc,
                              int* d_ncube, float* d_outerregion, int* d_sortpart,
                              tint2* d_beginendcube)
{
    GC_PARAMS_OR_RETURN(P);
    GC_REQUIRE(n >= 0, "CalFluidDen: n < 0");
    if (n == 0) return GC_OK;
    GC_REQUIRE(d_totalcoor && d_velrhop && d_cc && d_ncube && d_outerregion && d_sortpart &&
               d_beginendcube, "CalFluidDen: null pointer");
    k_density<<<blocks_for(n, BS), BS, 0, stream()>>>(n, P, d_totalcoor, d_velrhop, d_cc, d_ncube,
                                                     d_outerregion, d_sortpart, d_beginendcube);
    GC_LAUNCH_CHECK("CalFluidDen");
    return GC_OK;
}

extern "C" int gc_PredictFluidVel(int n, int* d_sortpart, tfloat4* d_velrhop,
                                  tfloat3* d_velhalf, tfloat3* d_totalforce)
{
    GC_PARAMS_OR_RETURN(P);
    GC_REQUIRE(n >= 0, "PredictFluidVel: n < 0");
    if (n == 0) return GC_OK;
    GC_REQUIRE(d_sortpart && d_velrhop && d_velhalf && d_totalforce, "PredictFluidVel: null pointer");
    k_predict<<<blocks_for(n, BS), BS, 0, stream()>>>(n, P, d_sortpart, d_velrhop, d_velhalf,
                                                     d_totalforce);
    GC_LAUNCH_CHECK("PredictFluidVel");
    return GC_OK;
}

extern "C" int gc_CalBound2Vel(int n, tfloat3* d_totalcoor, tfloat4* d_velrhop, tint3* d_cc,
                               int* d_ncube, float* d_outerregion, int* d_sortpart,
                               tint2* d_beginendcube)
{
    GC_PARAMS_OR_RETURN(P);
    GC_REQUIRE(n >= 0, "CalBound2Vel: n < 0");
    if (n == 0) return GC_OK;
    GC_REQUIRE(d_totalcoor && d_velrhop && d_cc && d_ncube && d_outerregion && d_sortpart &&
               d_beginendcube, "CalBound2Vel: null pointer");
    k_bound2_vel<<<blocks_for(n, BS), BS, 0, stream()>>>(n, P, d_totalcoor, d_velrhop, d_cc, d_ncube,
                                                        d_outerregion, d_sortpart, d_beginendcube);
    GC_LAUNCH_CHECK("CalBound2Vel");
    return GC_OK;
}

extern "C" int gc_EOS(int n, int* d_sortpart, tfloat4* d_velrhop, float* d_press)
{
    GC_PARAMS_OR_RETURN(P);
    GC_REQUIRE(n >= 0, "EOS: n < 0");
    if (n == 0) return GC_OK;
    GC_REQUIRE(d_sortpart && d_velrhop && d_press, "EOS: null pointer");
    k_eos<<<blocks_for(n, BS), BS, 0, stream()>>>(n, P, d_sortpart, d_velrhop, d_press);
    GC_LAUNCH_CHECK("EOS");
    return GC_OK;
}

extern "C" int gc_CalFluidSDPDForce(int n, tfloat3* h_sdpdforce, tfloat3* d_totalcoor,
                                    tfloat4* d_velrhop, tint3* d_cc, int* d_ncube,
                                    int* d_sortpart, float* d_outerreigon,
                                    tint2* d_beginendcube, tfloat3* d_totalforce,
                                    float* d_press)
{
    GC_PARAMS_OR_RETURN(P);
    GC_REQUIRE(n >= 0, "CalFluidSDPDForce: n < 0");
    if (n == 0) { advance_noise_counter(); return GC_OK; }
    GC_REQUIRE(d_totalcoor && d_velrhop && d_cc && d_ncube && d_sortpart && d_outerreigon &&
               d_beginendcube && d_totalforce && d_press, "CalFluidSDPDForce: null pointer");
    k_sdpd_force<<<blocks_for(n, BS), BS, 0, stream()>>>(n, P, d_totalcoor, d_velrhop, d_cc,
                                                        d_ncube, d_sortpart, d_outerreigon,
                                                        d_beginendcube, d_totalforce, d_press,
                                                        nullptr);
    GC_LAUNCH_CHECK("CalFluidSDPDForce");
    advance_noise_counter();
    if (h_sdpdforce) {
        int st = hip_check(hipMemcpyAsync(h_sdpdforce, d_totalforce, sizeof(tfloat3) * (size_t)n,
                                          hipMemcpyDeviceToHost, stream()), "SDPD force D2H");
        if (!st) st = hip_check(hipStreamSynchronize(stream()), "SDPD force sync");
        return st;
    }
    return GC_OK;
}

extern "C" int gc_CalFluidRepForce(int n, tfloat3* d_totalcoor, tint3* d_cc, int* d_ncube,
                                   int* d_sortpart, float* d_outerreigon,
                                   tint2* d_beginendcube, tfloat3* d_totalforce)
{
    GC_PARAMS_OR_RETURN(P);
    GC_REQUIRE(n >= 0, "CalFluidRepForce: n < 0");
    if (n == 0) return GC_OK;
    GC_REQUIRE(d_totalcoor && d_cc && d_ncube && d_sortpart && d_outerreigon && d_beginendcube &&
               d_totalforce, "CalFluidRepForce: null pointer");
    k_rep_force<<<blocks_for(n, BS), BS, 0, stream()>>>(n, P, d_totalcoor, d_cc, d_ncube, d_sortpart,
                                                       d_outerreigon, d_beginendcube, d_totalforce);
    GC_LAUNCH_CHECK("CalFluidRepForce");
    return GC_OK;
}

extern "C" int gc_SetFluidExtForce(int driveon, int n, int* d_sortpart, tfloat3* d_totalforce)
{
    GC_PARAMS_OR_RETURN(P);
    GC_REQUIRE(n >= 0, "SetFluidExtForce: n < 0");
    if (!driveon || n == 0) return GC_OK;
    if (P.force_mode != GC_FORCE_UNIFORM)
        return set_error(GC_ERR_INVALID,
                         "SetFluidExtForce: the reverse-Poiseuille pattern needs positions; "
                         "use gc_SetFluidExtForcePos");
    GC_REQUIRE(d_sortpart && d_totalforce, "SetFluidExtForce: null pointer");
    k_ext_force<<<blocks_for(n, BS), BS, 0, stream()>>>(n, P, d_sortpart, nullptr, d_totalforce);
    GC_LAUNCH_CHECK("SetFluidExtForce");
    return GC_OK;
}

extern "C" int gc_SetFluidExtForcePos(int driveon, int n, int* d_sortpart, tfloat3* d_totalcoor,
                                      tfloat3* d_totalforce)
{
    GC_PARAMS_OR_RETURN(P);
    GC_REQUIRE(n >= 0, "SetFluidExtForcePos: n < 0");
    if (!driveon || n == 0) return GC_OK;
    GC_REQUIRE(d_sortpart && d_totalcoor && d_totalforce, "SetFluidExtForcePos: null pointer");
    k_ext_force<<<blocks_for(n, BS), BS, 0, stream()>>>(n, P, d_sortpart, d_totalcoor, d_totalforce);
    GC_LAUNCH_CHECK("SetFluidExtForcePos");
    return GC_OK;
}

extern "C" int gc_CorrectFliudVel(int n, int* d_sortpart, tfloat3* d_totalforce,
                                  tfloat3* d_velhalf, tfloat4* d_velrhop, tfloat3* d_totalcoor,
                                  tint3* d_cc, int* d_ncube, float* d_outerreigon,
                                  tint2* d_beginendcube)
{
    (void)d_totalcoor; (void)d_cc; (void)d_ncube; (void)d_outerreigon; (void)d_beginendcube;
    GC_PARAMS_OR_RETURN(P);
    GC_REQUIRE(n >= 0, "CorrectFliudVel: n < 0");
    if (n == 0) return GC_OK;
    GC_REQUIRE(d_sortpart && d_totalforce && d_velhalf && d_velrhop, "CorrectFliudVel: null pointer");
    k_correct<<<blocks_for(n, BS), BS, 0, stream()>>>(n, P, d_sortpart, d_totalforce, d_velhalf,
                                                     d_velrhop);
    GC_LAUNCH_CHECK("CorrectFliudVel");
    return GC_OK;
}
