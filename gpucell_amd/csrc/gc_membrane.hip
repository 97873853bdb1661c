// gc_membrane.hip -- RBC membrane entry points replacing cupar::CalGeoPara,
// SetMechPara, CalCellDefForce and CalCellAggForce (src/ParticleGPU.cuh:25-35).
//
// Membrane vertex v (0-based) is particle nCellStart + v; its slot is
// curpos[nCellStart + v].  Triangle and edge records hold 1-based vertex ids
// (the reference's loader adds the per-cell offset to 1-based file ids,
// Particle.cu:252,268-284).  Unwrapped positions use the image counters:
// u = x + nthrobound * L on periodic axes; differences of unwrapped
// positions are formed in FP64 and rounded once.
//
// Mechanics (Fedosov, Caswell & Karniadakis 2010, Biophys J 98:2215):
//   E = sum_edges [ kT l_max/(4p) (3x^2-2x^3)/(1-x) + k_p/((m-1) l^(m-1))
//                   + k_b (1 - cos(theta - theta0)) ]
//     + sum_tri k_d (A_t - A_t0)^2/(2 A_t0)
//     + sum_cells [ k_a (A - A0)^2/(2 A0) + k_v (V - V0)^2/(2 V0) ],
// x = l/l_max, theta the signed dihedral of the edge's two triangles.
// Forces are -grad E, gathered per vertex over its incident edges and
// triangles in a fixed order (deterministic, no float atomics).
#include "gc_kernels.h"

#include <algorithm>
#include <vector>

using namespace gc;

namespace {

constexpr int BS = GC_BS;

struct Topo {
    const void* edge_ptr = nullptr;
    const void* tri_ptr = nullptr;
    int nedge = -1, ntri = -1, nm = -1;
    int* d_es = nullptr;   // [nm+1] edge-entry starts
    int* d_ee = nullptr;   // [4*nedge] (edge << 2) | role (0=A,1=B,2=C,3=D)
    int* d_ts = nullptr;   // [nm+1]
    int* d_te = nullptr;   // [3*ntri] (tri << 2) | corner
};
Topo g_topo;

void free_topo()
{
    if (g_topo.d_es) (void)hipFree(g_topo.d_es);
    if (g_topo.d_ee) (void)hipFree(g_topo.d_ee);
    if (g_topo.d_ts) (void)hipFree(g_topo.d_ts);
    if (g_topo.d_te) (void)hipFree(g_topo.d_te);
    g_topo = Topo();
}

}  // namespace

namespace gc {

// Builds vertex -> (edge, role) and vertex -> (triangle, corner) tables on
// the host from 1-based records (edge records may be null).  Entries of a
// vertex are in ascending element order.
int build_incidence(int nm, int nedge, const int6* h_edge, int ntri, const tint3* h_tri,
                    std::vector<int>& es, std::vector<int>& ee, std::vector<int>& ts,
                    std::vector<int>& te)
{
    es.assign(nm + 1, 0); ts.assign(nm + 1, 0);
    for (int e = 0; e < nedge; ++e) {
        const int* r = &h_edge[e].x;
        for (int k = 0; k < 4; ++k) {
            int v = r[k] - 1;
            if (v < 0 || v >= nm) return set_error(GC_ERR_TOPOLOGY, "edge vertex id out of range");
            es[v + 1]++;
        }
    }
    for (int t = 0; t < ntri; ++t) {
        const int* r = &h_tri[t].x;
        for (int k = 0; k < 3; ++k) {
            int v = r[k] - 1;
            if (v < 0 || v >= nm) return set_error(GC_ERR_TOPOLOGY, "triangle vertex id out of range");
            ts[v + 1]++;
        }
    }
    for (int v = 0; v < nm; ++v) { es[v + 1] += es[v]; ts[v + 1] += ts[v]; }
    ee.assign((size_t)4 * nedge, 0); te.assign((size_t)3 * ntri, 0);
    std::vector<int> fe(es.begin(), es.end() - 1), ft(ts.begin(), ts.end() - 1);
    for (int e = 0; e < nedge; ++e) {
        const int* r = &h_edge[e].x;
        for (int k = 0; k < 4; ++k) ee[fe[r[k] - 1]++] = (e << 2) | k;
    }
    for (int t = 0; t < ntri; ++t) {
        const int* r = &h_tri[t].x;
        for (int k = 0; k < 3; ++k) te[ft[r[k] - 1]++] = (t << 2) | k;
    }
    return GC_OK;
}

int ensure_topology(int nm, int nedge, const int6* d_edge, int ntri, const tint3* d_tri)
{
    if (g_topo.edge_ptr == (const void*)d_edge && g_topo.tri_ptr == (const void*)d_tri &&
        g_topo.nedge == nedge && g_topo.ntri == ntri && g_topo.nm == nm)
        return GC_OK;
    hipStream_t s = stream();
    int st = hip_check(hipStreamSynchronize(s), "topology sync");
    if (st) return st;
    std::vector<int6> he(d_edge ? nedge : 0);
    std::vector<tint3> ht(ntri);
    if (d_edge && nedge > 0)
        if ((st = hip_check(hipMemcpy(he.data(), d_edge, sizeof(int6) * nedge, hipMemcpyDeviceToHost), "edge D2H"))) return st;
    if (ntri > 0)
        if ((st = hip_check(hipMemcpy(ht.data(), d_tri, sizeof(tint3) * ntri, hipMemcpyDeviceToHost), "tri D2H"))) return st;
    std::vector<int> es, ee, ts, te;
    st = build_incidence(nm, d_edge ? nedge : 0, he.data(), ntri, ht.data(), es, ee, ts, te);
    if (st) return st;
    free_topo();
    auto up = [&](int** dst, const std::vector<int>& v) -> int {
        size_t b = sizeof(int) * (v.size() ? v.size() : 1);
        if (hipMalloc(dst, b) != hipSuccess) return set_error(GC_ERR_NOMEM, "topology alloc");
        if (!v.empty()) return hip_check(hipMemcpy(*dst, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice), "topology H2D");
        return GC_OK;
    };
    if ((st = up(&g_topo.d_es, es)) || (st = up(&g_topo.d_ee, ee)) || (st = up(&g_topo.d_ts, ts)) ||
        (st = up(&g_topo.d_te, te))) { free_topo(); return st; }
    g_topo.edge_ptr = d_edge; g_topo.tri_ptr = d_tri;
    g_topo.nedge = d_edge ? nedge : 0; g_topo.ntri = ntri; g_topo.nm = nm;
    return GC_OK;
}

struct PerBox { float L[3]; int per[3]; };

__device__ __forceinline__ PerBox load_box(const GcParams& P, const float* bbox)
{
    PerBox b;
    for (int d = 0; d < 3; ++d) { b.L[d] = bbox[2 * d + 1] - bbox[2 * d]; b.per[d] = P.per[d]; }
    return b;
}

// unwrapped(u) - unwrapped(o) in FP64, rounded to FP32
__device__ __forceinline__ float3 urel(const tfloat3& xu, const tint3& nu, const tfloat3& xo,
                                       const tint3& no, const PerBox& b)
{
    double dx = (double)xu.x - (double)xo.x, dy = (double)xu.y - (double)xo.y,
           dz = (double)xu.z - (double)xo.z;
    if (b.per[0]) dx += (double)(nu.x - no.x) * (double)b.L[0];
    if (b.per[1]) dy += (double)(nu.y - no.y) * (double)b.L[1];
    if (b.per[2]) dz += (double)(nu.z - no.z) * (double)b.L[2];
    return make_float3((float)dx, (float)dy, (float)dz);
}

__device__ __forceinline__ float3 f3sub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 f3cross(float3 a, float3 b)
{
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float f3dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 f3scale(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }

// ---- CalGeoPara ---------------------------------------------------------
// One block per cell: per-triangle geometry, FP64 block reduction of the
// cell's area and volume (signed tetrahedra against the cell's first
// triangle's first vertex), then the per-triangle broadcasts.
__global__ __launch_bounds__(GC_GEO_BS) void k_geo_cell(GcParams P, long stepcount,
    const GcRep* __restrict__ reps, const int* __restrict__ cell_rep,
    const float* __restrict__ bbox, const int* __restrict__ curpos, const tint3* __restrict__ tri,
    const tint2* __restrict__ betri, const tint3* __restrict__ nthro, const tfloat3* __restrict__ coor,
    tfloat3* __restrict__ trinor, float* __restrict__ triarea, float* __restrict__ reftriarea,
    tfloat3* __restrict__ tricen, float* __restrict__ refarea_para, float* __restrict__ refvol_para,
    float* __restrict__ area_para, float* __restrict__ vol_para, float* __restrict__ area,
    float* __restrict__ vol)
{
    constexpr int BS = GC_GEO_BS;
    __shared__ double sa[BS], sv[BS];
    const int c = blockIdx.x;
    const int tb = betri[c].x, te = betri[c].y;
    if (reps) {  // batched: this cell's suspension
        const GcRep& R = reps[cell_rep[c]];
        P = R.P;
        bbox = R.bbox;
    }
    PerBox b = load_box(P, bbox);
    int so = curpos[P.nc + tri[tb].x - 1];
    tfloat3 xo = coor[so];
    tint3 no = nthro[so];
    double A = 0.0, V = 0.0;
    for (int t = tb + threadIdx.x; t <= te; t += BS) {
        tint3 tr = tri[t];
        int s0 = curpos[P.nc + tr.x - 1], s1 = curpos[P.nc + tr.y - 1], s2 = curpos[P.nc + tr.z - 1];
        float3 p0 = urel(coor[s0], nthro[s0], xo, no, b);
        float3 p1 = urel(coor[s1], nthro[s1], xo, no, b);
        float3 p2 = urel(coor[s2], nthro[s2], xo, no, b);
        float3 n = f3cross(f3sub(p1, p0), f3sub(p2, p0));
        const float n2 = f3dot(n, n);
        float in = __builtin_amdgcn_rsqf(n2);
        float at = n2 > 0.0f ? 0.5f * (n2 * in) : 0.0f;   // area 0 for a degenerate triangle
        trinor[t] = tfloat3{n.x * in, n.y * in, n.z * in};
        triarea[t] = at;
        if (stepcount == -1) reftriarea[t] = at;
        // centre in absolute unwrapped coordinates of the cell's frame
        double ox = (double)xo.x + (b.per[0] ? (double)no.x * b.L[0] : 0.0);
        double oy = (double)xo.y + (b.per[1] ? (double)no.y * b.L[1] : 0.0);
        double oz = (double)xo.z + (b.per[2] ? (double)no.z * b.L[2] : 0.0);
        constexpr double third = 1.0 / 3.0;
        tricen[t] = tfloat3{(float)(ox + ((double)p0.x + p1.x + p2.x) * third),
                            (float)(oy + ((double)p0.y + p1.y + p2.y) * third),
                            (float)(oz + ((double)p0.z + p1.z + p2.z) * third)};
        A += (double)at;
        // signed tetrahedron (O, p0, p1, p2) in FP64: six times its volume
        double ax = p0.x, ay = p0.y, az = p0.z, bx = p1.x, by = p1.y, bz = p1.z;
        double cx = p2.x, cy = p2.y, cz = p2.z;
        V += ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    }
    sa[threadIdx.x] = A; sv[threadIdx.x] = V;
    __syncthreads();
    for (int o = BS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) { sa[threadIdx.x] += sa[threadIdx.x + o]; sv[threadIdx.x] += sv[threadIdx.x + o]; }
        __syncthreads();
    }
    float Ac = (float)sa[0], Vc = (float)(sv[0] / 6.0);
    if (threadIdx.x == 0) { area[c] = Ac; vol[c] = Vc; }
    for (int t = tb + threadIdx.x; t <= te; t += BS) {
        area_para[t] = Ac; vol_para[t] = Vc;
        if (stepcount == -1) { refarea_para[t] = Ac; refvol_para[t] = Vc; }
    }
}

__global__ __launch_bounds__(BS) void k_par_area(int nm, const int* __restrict__ ts,
                                                 const int* __restrict__ te,
                                                 const float* __restrict__ triarea,
                                                 float* __restrict__ pararea)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nm) return;
    float a = 0.f;
    for (int q = ts[v]; q < ts[v + 1]; ++q) a += triarea[te[q] >> 2];
    pararea[v] = a * (1.0f / 3.0f);
}

// ---- SetMechPara ----------------------------------------------------------
// Dihedral of edge (A,B) with wings C (triangle v) and D (triangle w), all
// relative to A: n1 = e x c, n2 = d x e, theta = atan2((n1^ x n2^).e^, n1^.n2^).
__device__ __forceinline__ double dihedral_d(double3 e, double3 c, double3 d)
{
    double3 n1 = make_double3(e.y * c.z - e.z * c.y, e.z * c.x - e.x * c.z, e.x * c.y - e.y * c.x);
    double3 n2 = make_double3(d.y * e.z - d.z * e.y, d.z * e.x - d.x * e.z, d.x * e.y - d.y * e.x);
    double m1 = sqrt(n1.x * n1.x + n1.y * n1.y + n1.z * n1.z);
    double m2 = sqrt(n2.x * n2.x + n2.y * n2.y + n2.z * n2.z);
    double l = sqrt(e.x * e.x + e.y * e.y + e.z * e.z);
    double3 u1 = make_double3(n1.x / m1, n1.y / m1, n1.z / m1);
    double3 u2 = make_double3(n2.x / m2, n2.y / m2, n2.z / m2);
    double3 x = make_double3(u1.y * u2.z - u1.z * u2.y, u1.z * u2.x - u1.x * u2.z, u1.x * u2.y - u1.y * u2.x);
    return atan2((x.x * e.x + x.y * e.y + x.z * e.z) / l, u1.x * u2.x + u1.y * u2.y + u1.z * u2.z);
}

__global__ __launch_bounds__(BS) void k_mech_para(GcParams P, int nedge, const int6* __restrict__ edge,
    const int* __restrict__ curpos, const tfloat3* __restrict__ coor, float* __restrict__ l0o,
    float* __restrict__ lmaxo, const float* __restrict__ kc, const float* __restrict__ mu0,
    float* __restrict__ kbo, float* __restrict__ mo, float* __restrict__ po, float* __restrict__ kpo,
    float* __restrict__ th0o)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nedge) return;
    int6 E = edge[e];
    tfloat3 A = coor[curpos[P.nc + E.x - 1]], B = coor[curpos[P.nc + E.y - 1]];
    tfloat3 C = coor[curpos[P.nc + E.z - 1]], D = coor[curpos[P.nc + E.k - 1]];
    double3 be = make_double3((double)B.x - A.x, (double)B.y - A.y, (double)B.z - A.z);
    double3 ce = make_double3((double)C.x - A.x, (double)C.y - A.y, (double)C.z - A.z);
    double3 de = make_double3((double)D.x - A.x, (double)D.y - A.y, (double)D.z - A.z);
    double L0 = sqrt(be.x * be.x + be.y * be.y + be.z * be.z);
    double x0 = 1.0 / (double)P.wlc_ratio, m = P.pow_index, kT = P.kT;
    double g = 1.0 / (4.0 * (1 - x0) * (1 - x0)) - 0.25 + x0;
    double hh = x0 / (2.0 * (1 - x0) * (1 - x0) * (1 - x0)) - 1.0 / (4.0 * (1 - x0) * (1 - x0)) + 0.25;
    double p = kT * sqrt(3.0) * (hh + (m + 1.0) * g) / (4.0 * (double)mu0[e] * L0);
    l0o[e] = (float)L0;
    lmaxo[e] = (float)(L0 * (double)P.wlc_ratio);
    mo[e] = (float)m;
    po[e] = (float)p;
    kpo[e] = (float)(kT / p * g * pow(L0, m));
    kbo[e] = (float)(2.0 * (double)kc[e] / sqrt(3.0));
    th0o[e] = (float)dihedral_d(be, ce, de);
}

// ---- CalCellDefForce --------------------------------------------------------
// Position of vertex v1 (1-based) relative to this vertex: an FP64
// difference of unwrapped positions, rounded once (both paths give the same
// bits: x + n L is exact in FP64 for an FP32 x).
__device__ __forceinline__ float3 vpos(const DefArgs& a, const GcParams& P, const PerBox& b, int v1,
                                       const tfloat3& xo, const tint3& no, double ox, double oy,
                                       double oz)
{
    if (a.mpos) {
        int u = v1 - 1;
        return make_float3((float)(a.mpos[u] - ox), (float)(a.mpos[a.nm + u] - oy),
                           (float)(a.mpos[2 * a.nm + u] - oz));
    }
    int s = a.curpos[P.nc + v1 - 1];
    return urel(a.coor[s], a.nthro[s], xo, no, b);
}

__global__ __launch_bounds__(BS) void k_def_force(GcParams P, int nm, const float* __restrict__ bbox,
                                                  DefArgs a)
{
    int v = xcd_block((nm + (int)blockDim.x - 1) / (int)blockDim.x) * blockDim.x + threadIdx.x;
    if (v >= nm) return;
    if (a.reps) {  // batched: the vertex's suspension
        const GcRep& R = a.reps[a.cell_rep[a.celltab[v]]];
        P = R.P;
        bbox = R.bbox;
    }
    PerBox b = load_box(P, bbox);
    int sv = a.curpos[P.nc + v];
    tfloat3 xo = a.coor[sv];
    tint3 no = a.nthro[sv];
    double ox = 0.0, oy = 0.0, oz = 0.0;
    if (a.mpos) { ox = a.mpos[v]; oy = a.mpos[nm + v]; oz = a.mpos[2 * nm + v]; }
    float fx = 0.f, fy = 0.f, fz = 0.f;
    // edges: WLC-POW spring (roles A, B) and bending (all four roles)
    for (int q = a.es[v]; q < a.es[v + 1]; ++q) {
        int code = a.ee[q], e = code >> 2, role = code & 3;
        int6 E = a.edge[e];
        float3 pA = vpos(a, P, b, E.x, xo, no, ox, oy, oz), pB = vpos(a, P, b, E.y, xo, no, ox, oy, oz);
        float3 pC = vpos(a, P, b, E.z, xo, no, ox, oy, oz), pD = vpos(a, P, b, E.k, xo, no, ox, oy, oz);
        if (role < 2) {
            float3 d = f3sub(pA, pB);
            float l = sqrtf(f3dot(d, d));
            float x = l / a.lmax[e];
            float om = 1.0f - x;
            float fm = -(P.kT / a.p[e]) * (0.25f / (om * om) - 0.25f + x) + a.kp[e] / powf(l, a.mexp[e]);
            float s = (role == 0 ? fm : -fm) / l;
            fx += s * d.x; fy += s * d.y; fz += s * d.z;
        }
        float3 e3 = f3sub(pB, pA), ac = f3sub(pC, pA), ad = f3sub(pD, pA);
        float3 n1 = f3cross(e3, ac), n2 = f3cross(ad, e3);
        float l2 = f3dot(e3, e3), l = sqrtf(l2);
        float q1 = f3dot(n1, n1), q2 = f3dot(n2, n2);
        float im1 = rsqrtf(q1), im2 = rsqrtf(q2);
        float3 u1 = f3scale(n1, im1), u2 = f3scale(n2, im2);
        float3 xx = f3cross(u1, u2);
        float th = atan2f(f3dot(xx, e3) / l, f3dot(u1, u2));
        float cb = -a.kb[e] * sinf(th - a.th0[e]);
        float3 gC = f3scale(n1, -l / q1), gD = f3scale(n2, -l / q2);
        float3 g;
        if (role == 2) g = gC;
        else if (role == 3) g = gD;
        else {
            float tC = f3dot(ac, e3) / l2, tD = f3dot(ad, e3) / l2;
            float wC = role == 0 ? -(1.0f - tC) : -tC;
            float wD = role == 0 ? -(1.0f - tD) : -tD;
            g = make_float3(wC * gC.x + wD * gD.x, wC * gC.y + wD * gD.y, wC * gC.z + wD * gD.z);
        }
        fx += cb * g.x; fy += cb * g.y; fz += cb * g.z;
    }
    // triangles: local + global area, volume (tetrahedra against this vertex)
    for (int q = a.ts[v]; q < a.ts[v + 1]; ++q) {
        int code = a.te[q], t = code >> 2, k = code & 3;
        tint3 T = a.tri[t];
        int id[3] = {T.x, T.y, T.z};
        float3 P0 = vpos(a, P, b, id[0], xo, no, ox, oy, oz), P1 = vpos(a, P, b, id[1], xo, no, ox, oy, oz);
        float3 P2 = vpos(a, P, b, id[2], xo, no, ox, oy, oz);
        float3 n = f3cross(f3sub(P1, P0), f3sub(P2, P0));
        float nn = sqrtf(f3dot(n, n));
        float At = 0.5f * nn;
        float3 un = f3scale(n, 1.0f / nn);
        float cA = a.ka[t] * (a.Ac[t] - a.A0c[t]) / a.A0c[t] + a.kd[t] * (At - a.reftriarea[t]) / a.reftriarea[t];
        float cV = a.kv[t] * (a.Vc[t] - a.V0c[t]) / a.V0c[t];
        float3 pn = k == 0 ? P1 : (k == 1 ? P2 : P0);   // next corner
        float3 pnn = k == 0 ? P2 : (k == 1 ? P0 : P1);  // next-next corner
        float3 dA = f3cross(un, f3sub(pnn, pn));        // 2 dA/dP_k
        float3 dV = f3cross(pn, pnn);                   // 6 dV/dP_k (this vertex at origin)
        fx += -0.5f * cA * dA.x - cV * dV.x * (1.0f / 6.0f);
        fy += -0.5f * cA * dA.y - cV * dV.y * (1.0f / 6.0f);
        fz += -0.5f * cA * dA.z - cV * dV.z * (1.0f / 6.0f);
    }
    a.defforce[v] = tfloat3{fx, fy, fz};
    if (!a.force) return;   // membrane-only evaluation (substepping)
    tfloat3 F = a.force[sv];
    F.x += fx; F.y += fy; F.z += fz;
    a.force[sv] = F;
    if (a.vout) {  // engine: this is the vertex's last force term -> corrector
        float h = 0.5f * P.dt * P.inv_m;
        tfloat3 vh = a.velhalf[sv];
        a.vout[sv] = make_float4(vh.x + h * F.x, vh.y + h * F.y, vh.z + h * F.z, a.velrho[sv].w);
    }
}

// ---- CalCellDefForce in three passes (engine) -----------------------------
// The gather above evaluates every edge's spring and dihedral from each of
// its four vertices and every triangle from each of its three corners.  The
// engine evaluates each term once instead: one thread per edge writes the
// edge's forces on its four vertices (roles A, B, C, D), one thread per
// triangle the area/volume forces on its three corners, and one thread per
// vertex sums its entries in the incidence order above.  Positions are
// relative to the edge's (triangle's) first vertex; every term depends on
// differences only.  The volume gradient of a triangle at corner k with
// that corner as origin is (P_n - P_k) x (P_nn - P_k), the triangle's
// un-normalised normal n = (P1 - P0) x (P2 - P0) for every corner.

__device__ __forceinline__ void rep_params(const DefArgs& a, int vtx0, GcParams& P, const float*& bbox)
{
    if (a.reps) {
        const GcRep& R = a.reps[a.cell_rep[a.celltab[vtx0]]];
        P = R.P;
        bbox = R.bbox;
    }
}

// origin of a record: its first vertex (1-based id v1)
__device__ __forceinline__ void rec_origin(const DefArgs& a, const GcParams& P, int v1, tfloat3& xo,
                                           tint3& no, double& ox, double& oy, double& oz)
{
    if (a.mpos) {
        const int u = v1 - 1;
        ox = a.mpos[u]; oy = a.mpos[a.nm + u]; oz = a.mpos[2 * a.nm + u];
        xo = tfloat3{0.f, 0.f, 0.f}; no = tint3{0, 0, 0};
    } else {
        const int s = a.curpos[P.nc + v1 - 1];
        xo = a.coor[s]; no = a.nthro[s];
        ox = oy = oz = 0.0;
    }
}

// per-edge and per-triangle parameters (loaded ahead of their use)
struct EdgePar { float lmax, p, kp, mexp, kb, cth0, sth0; };
struct TriPar { float ka, A0c, kd, ref, kv, V0c; };
__device__ __forceinline__ EdgePar edge_par(const DefArgs& a, int e)
{
    float c, sn;
    if (a.cth0) {   // precomputed (engine)
        c = a.cth0[e];
        sn = a.sth0[e];
    } else {
        sincosf(a.th0[e], &sn, &c);
    }
    return EdgePar{a.lmax[e], a.p[e], a.kp[e], a.mexp[e], a.kb[e], c, sn};
}
__device__ __forceinline__ TriPar tri_par(const DefArgs& a, int t)
{
    return TriPar{a.ka[t], a.A0c[t], a.kd[t], a.reftriarea[t], a.kv[t], a.V0c[t]};
}

// Forces of an edge on its vertices A, B, C, D from B, C, D relative to A.
// Quotients and roots below use the hardware reciprocal, reciprocal square
// root, base-2 logarithm and exponential (v_rcp/v_rsq/v_log/v_exp, ~1 ulp):
// the terms stay well inside the 1e-5 force tolerance at a fraction of the
// instructions of IEEE division and the library functions.
__device__ __forceinline__ float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq_hw(float x) { return __builtin_amdgcn_rsqf(x); }

// bad: GC_FAULT_WLC is OR-ed in when the edge is at or past l_max (the
// WLC force is singular at x = 1 and has the wrong sign beyond it).
__device__ __forceinline__ void edge_terms(const GcParams& P, const EdgePar& q, float3 pB, float3 pC, float3 pD,
                                           float3 out[4], int& bad)
{
    // WLC-POW spring along d = A - B: s d on A, -s d on B
    const float3 d = make_float3(-pB.x, -pB.y, -pB.z);
    const float dd = f3dot(d, d);
    const float rls = rsq_hw(dd);
    const float ls = dd * rls;
    const float x = ls * rcp_hw(q.lmax);
    bad |= x >= 1.0f ? GC_FAULT_WLC : 0;
    const float iom = rcp_hw(1.0f - x);
    // kp / ls^m = kp 2^(-m log2 ls)
    const float fm = -(P.kT * rcp_hw(q.p)) * (0.25f * iom * iom - 0.25f + x) +
                     q.kp * __builtin_amdgcn_exp2f(-q.mexp * __builtin_amdgcn_logf(ls));
    const float s = fm * rls;
    // bending: dihedral of the wings C (n1 = e x c) and D (n2 = d x e)
    const float3 e3 = pB, ac = pC, ad = pD;
    const float3 n1 = f3cross(e3, ac), n2 = f3cross(ad, e3);
    const float l2 = f3dot(e3, e3), rl = rsq_hw(l2), l = l2 * rl;
    const float q1 = f3dot(n1, n1), q2 = f3dot(n2, n2);
    const float im1 = rsq_hw(q1), im2 = rsq_hw(q2);
    const float3 u1 = f3scale(n1, im1), u2 = f3scale(n2, im2);
    // dihedral angle th from (cos th, sin th) = (u1.u2, (u1 x u2).e3 / l) / r,
    // sin(th - th0) = sin th cos th0 - cos th sin th0
    const float3 xx = f3cross(u1, u2);
    const float sy = f3dot(xx, e3) * rl, cx = f3dot(u1, u2);
    const float ir = rsq_hw(cx * cx + sy * sy);
    const float cb = -q.kb * ((sy * q.cth0 - cx * q.sth0) * ir);
    const float3 gC = f3scale(n1, -l * (im1 * im1)), gD = f3scale(n2, -l * (im2 * im2));
    const float rl2 = rl * rl;
    const float tC = f3dot(ac, e3) * rl2, tD = f3dot(ad, e3) * rl2;
    const float wCA = -(1.0f - tC), wDA = -(1.0f - tD), wCB = -tC, wDB = -tD;
    out[0] = make_float3(s * d.x + cb * (wCA * gC.x + wDA * gD.x), s * d.y + cb * (wCA * gC.y + wDA * gD.y),
                         s * d.z + cb * (wCA * gC.z + wDA * gD.z));
    out[1] = make_float3(-s * d.x + cb * (wCB * gC.x + wDB * gD.x), -s * d.y + cb * (wCB * gC.y + wDB * gD.y),
                         -s * d.z + cb * (wCB * gC.z + wDB * gD.z));
    out[2] = make_float3(cb * gC.x, cb * gC.y, cb * gC.z);
    out[3] = make_float3(cb * gD.x, cb * gD.y, cb * gD.z);
}

// Area and volume forces of triangle t on its corners from P1, P2 relative
// to P0; Ac, Vc the cell's current area and volume.
__device__ __forceinline__ void tri_terms(const TriPar& q, float3 P1, float3 P2, float Ac, float Vc, float3 out[3])
{
    const float3 P0 = make_float3(0.f, 0.f, 0.f);
    const float3 n = f3cross(P1, P2);
    const float n2 = f3dot(n, n);
    const float inn = rsq_hw(n2);
    const float At = 0.5f * (n2 * inn);
    const float3 un = f3scale(n, inn);
    const float cA = q.ka * (Ac - q.A0c) * rcp_hw(q.A0c) + q.kd * (At - q.ref) * rcp_hw(q.ref);
    const float cV = q.kv * (Vc - q.V0c) * rcp_hw(q.V0c);
    const float hv = cV * (1.0f / 6.0f);
    const float3 cn[3] = {P1, P2, P0};   // next corner of corner k
    const float3 cnn[3] = {P2, P0, P1};  // next-next corner
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float3 dA = f3cross(un, f3sub(cnn[k], cn[k]));   // 2 dA/dP_k
        out[k] = make_float3(-0.5f * cA * dA.x - hv * n.x, -0.5f * cA * dA.y - hv * n.y,
                             -0.5f * cA * dA.z - hv * n.z);
    }
}

__global__ __launch_bounds__(BS) void k_edge_force(GcParams P, int nedge, const float* __restrict__ bbox,
                                                   DefArgs a, tfloat3* __restrict__ ef)
{
    const int e = xcd_block((nedge + (int)blockDim.x - 1) / (int)blockDim.x) * blockDim.x + threadIdx.x;
    if (e >= nedge) return;
    const int6 E = a.edge[e];
    const float* bb = bbox;
    rep_params(a, E.x - 1, P, bb);
    const PerBox b = load_box(P, bb);
    tfloat3 xo; tint3 no; double ox, oy, oz;
    rec_origin(a, P, E.x, xo, no, ox, oy, oz);
    const float3 pB = vpos(a, P, b, E.y, xo, no, ox, oy, oz);
    const float3 pC = vpos(a, P, b, E.z, xo, no, ox, oy, oz);
    const float3 pD = vpos(a, P, b, E.k, xo, no, ox, oy, oz);
    float3 f[4];
    int bad = 0;
    edge_terms(P, edge_par(a, e), pB, pC, pD, f, bad);
    if (bad && a.fault) atomicOr(a.fault, bad);
    tfloat3* o = ef + 4 * (size_t)e;
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = tfloat3{f[k].x, f[k].y, f[k].z};
}

__global__ __launch_bounds__(BS) void k_tri_force(GcParams P, int ntri, const float* __restrict__ bbox,
                                                  DefArgs a, tfloat3* __restrict__ tf)
{
    const int t = xcd_block((ntri + (int)blockDim.x - 1) / (int)blockDim.x) * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    const tint3 T = a.tri[t];
    const float* bb = bbox;
    rep_params(a, T.x - 1, P, bb);
    const PerBox b = load_box(P, bb);
    tfloat3 xo; tint3 no; double ox, oy, oz;
    rec_origin(a, P, T.x, xo, no, ox, oy, oz);
    const float3 P1 = vpos(a, P, b, T.y, xo, no, ox, oy, oz);
    const float3 P2 = vpos(a, P, b, T.z, xo, no, ox, oy, oz);
    float3 f[3];
    tri_terms(tri_par(a, t), P1, P2, a.Ac[t], a.Vc[t], f);
    tfloat3* o = tf + 3 * (size_t)t;
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = tfloat3{f[k].x, f[k].y, f[k].z};
}

__global__ __launch_bounds__(BS) void k_vert_force(GcParams P, int nm, DefArgs a,
                                                   const tfloat3* __restrict__ ef,
                                                   const tfloat3* __restrict__ tf)
{
    const int v = xcd_block((nm + (int)blockDim.x - 1) / (int)blockDim.x) * blockDim.x + threadIdx.x;
    if (v >= nm) return;
    if (a.reps) P = a.reps[a.cell_rep[a.celltab[v]]].P;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    for (int q = a.es[v]; q < a.es[v + 1]; ++q) {   // code = 4 edge + role
        const tfloat3 g = ef[a.ee[q]];
        fx += g.x; fy += g.y; fz += g.z;
    }
    for (int q = a.ts[v]; q < a.ts[v + 1]; ++q) {
        const int code = a.te[q];
        const tfloat3 g = tf[3 * (code >> 2) + (code & 3)];
        fx += g.x; fy += g.y; fz += g.z;
    }
    if (a.fault && !isfinite(fx + fy + fz)) atomicOr(a.fault, GC_FAULT_NONFINITE);
    a.defforce[v] = tfloat3{fx, fy, fz};
    if (!a.force) return;   // membrane-only evaluation (substepping)
    const int sv = a.curpos[P.nc + v];
    tfloat3 F = a.force[sv];
    F.x += fx; F.y += fy; F.z += fz;
    a.force[sv] = F;
    if (a.vout) {  // this is the vertex's last force term -> corrector
        const float h = 0.5f * P.dt * P.inv_m;
        const tfloat3 vh = a.velhalf[sv];
        a.vout[sv] = make_float4(vh.x + h * F.x, vh.y + h * F.y, vh.z + h * F.z, a.velrho[sv].w);
    }
}

// A vertex's incidence list [q0, q1) of (record << 2 | role) codes: the
// first NC codes loaded together from clamped indices, and the sum of the
// matching LDS terms (w floats per record) in list order.
template <int NC>
struct IncList {
    int q0 = 0, q1 = 0;
    int cd[NC];
    __device__ __forceinline__ void load(const int* __restrict__ off, const int* __restrict__ code, int g)
    {
        q0 = off[g];
        q1 = off[g + 1];
#pragma unroll
        for (int k = 0; k < NC; ++k) cd[k] = code[max(min(q0 + k, q1 - 1), 0)];
    }
    __device__ __forceinline__ void sum(const float* buf, int w, int base, const int* __restrict__ code, float& fx,
                                        float& fy, float& fz) const
    {
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            if (q0 + k >= q1) break;
            const float* p = buf + w * ((cd[k] >> 2) - base) + 3 * (cd[k] & 3);
            fx += p[0]; fy += p[1]; fz += p[2];
        }
        for (int q = q0 + NC; q < q1; ++q) {   // more incidences than NC: one by one
            const int c = code[q];
            const float* p = buf + w * ((c >> 2) - base) + 3 * (c & 3);
            fx += p[0]; fy += p[1]; fz += p[2];
        }
    }
};
using EdgeInc = IncList<GC_CELL_INC>;
using TriInc = IncList<GC_CELL_TINC>;

// Geometry and membrane force of one cell from the FP64 unwrapped positions
// of its vertices in sx/sy/sz (LDS, local index); on return thread tid < nv
// holds the force on vertex v0 + tid in f.  With geo the geometry outputs
// (normals, areas, centroids, cell area and volume) are written.  Ends after
// reading buf: a caller that refills sx or buf first passes a barrier.
// With PRE the vertex's incidence lists (ei, ti) were loaded by the caller
// (ahead of the phases that need them); otherwise they are loaded here.
template <bool PRE>
__device__ __forceinline__ void cell_force(const GcParams& P, const CellArgs& ca, const DefArgs& a, int c, int tid,
                                           int v0, int nv, int e0, int ne, int tb, int te, bool geo,
                                           const double* sx, const double* sy, const double* sz, float* buf,
                                           double* sa, double* sv, float3& f, const EdgeInc& ei_pre,
                                           const TriInc& ti_pre)
{
    constexpr int BS = GC_CELL_BS;
    // differences of unwrapped positions rounded once (what urel and vpos
    // compute); the substepper's geometry input is the FP32 rounding of them
    auto rel = [&](int u, int o) {
        return make_float3((float)(sx[u] - sx[o]), (float)(sy[u] - sy[o]), (float)(sz[u] - sz[o]));
    };
    auto rel_geo = [&](int u, int o) {
        if (!ca.round_geo) return rel(u, o);
        return make_float3((float)((double)(float)sx[u] - (double)(float)sx[o]),
                           (float)((double)(float)sy[u] - (double)(float)sy[o]),
                           (float)((double)(float)sz[u] - (double)(float)sz[o]));
    };
    // geometry: k_geo_cell's arithmetic and reduction order (GC_GEO_BS =
    // GC_CELL_BS strided sums, then a tree).
    // Global loads are issued in groups from clamped indices before their
    // results are used (a load under a branch would wait out its latency
    // alone); the arithmetic and its order are unchanged.
    const int o = a.tri[tb].x - 1 - v0;
    double A = 0.0, V = 0.0;
    static_assert(GC_GEO_BS == GC_CELL_BS, "geometry reduction order");
    {
        const double ox = ca.round_geo ? (double)(float)sx[o] : sx[o];
        const double oy = ca.round_geo ? (double)(float)sy[o] : sy[o];
        const double oz = ca.round_geo ? (double)(float)sz[o] : sz[o];
        for (int t0 = tb + tid; t0 <= te; t0 += 2 * BS) {
        tint3 trs[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) trs[u] = a.tri[min(t0 + BS * u, te)];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int t = t0 + BS * u;
            if (t > te) break;
            const tint3 tr = trs[u];
            const float3 p0 = rel_geo(tr.x - 1 - v0, o), p1 = rel_geo(tr.y - 1 - v0, o),
                         p2 = rel_geo(tr.z - 1 - v0, o);
            const float3 n = f3cross(f3sub(p1, p0), f3sub(p2, p0));
            const float n2 = f3dot(n, n);
            const float in = __builtin_amdgcn_rsqf(n2);
            const float at = n2 > 0.0f ? 0.5f * (n2 * in) : 0.0f;
            if (geo) {
                ca.trinor[t] = tfloat3{n.x * in, n.y * in, n.z * in};
                ca.triarea[t] = at;
                if (ca.step == -1) ca.reftriarea[t] = at;
                constexpr double third = 1.0 / 3.0;
                ca.tricen[t] = tfloat3{(float)(ox + ((double)p0.x + p1.x + p2.x) * third),
                                       (float)(oy + ((double)p0.y + p1.y + p2.y) * third),
                                       (float)(oz + ((double)p0.z + p1.z + p2.z) * third)};
            }
            A += (double)at;
            double ax = p0.x, ay = p0.y, az = p0.z, bx = p1.x, by = p1.y, bz = p1.z;
            double cx = p2.x, cy = p2.y, cz = p2.z;
            V += ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
        }
        }
        sa[tid] = A; sv[tid] = V;
    }
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (tid < w) { sa[tid] += sa[tid + w]; sv[tid] += sv[tid + w]; }
        __syncthreads();
    }
    const float Ac = (float)sa[0], Vc = (float)(sv[0] / 6.0);
    if (geo) {
        if (tid == 0) { ca.area[c] = Ac; ca.vol[c] = Vc; }
        for (int t = tb + tid; t <= te; t += BS) {
            ca.area_para[t] = Ac; ca.vol_para[t] = Vc;
            if (ca.step == -1) { ca.refarea_para[t] = Ac; ca.refvol_para[t] = Vc; }
        }
    }
    // edges -> buf[12 (e - e0) + 3 role]
    int bad = 0;
    for (int i0 = tid; i0 < ne; i0 += 2 * BS) {
        int6 Es[2];
        EdgePar qs[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = e0 + min(i0 + BS * u, ne - 1);
            Es[u] = a.edge[e];
            qs[u] = edge_par(a, e);
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int i = i0 + BS * u;
            if (i >= ne) break;
            const int6 E = Es[u];
            const int ua = E.x - 1 - v0;
            float3 fe[4];
            edge_terms(P, qs[u], rel(E.y - 1 - v0, ua), rel(E.z - 1 - v0, ua), rel(E.k - 1 - v0, ua), fe, bad);
            float* q = buf + 12 * i;
#pragma unroll
            for (int k = 0; k < 4; ++k) { q[3 * k] = fe[k].x; q[3 * k + 1] = fe[k].y; q[3 * k + 2] = fe[k].z; }
        }
    }
    __syncthreads();
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (tid < nv) {
        if constexpr (PRE) {
            ei_pre.sum(buf, 12, e0, a.ee, fx, fy, fz);
        } else {
            EdgeInc ei;
            ei.load(a.es, a.ee, v0 + tid);
            ei.sum(buf, 12, e0, a.ee, fx, fy, fz);
        }
    }
    __syncthreads();
    // triangles -> buf[9 (t - tb) + 3 corner]
    const int nt = te - tb + 1;
    for (int i0 = tid; i0 < nt; i0 += 2 * BS) {
        tint3 Ts[2];
        TriPar qs[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int t = tb + min(i0 + BS * u, nt - 1);
            Ts[u] = a.tri[t];
            qs[u] = tri_par(a, t);
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int i = i0 + BS * u;
            if (i >= nt) break;
            const tint3 T = Ts[u];
            const int u0 = T.x - 1 - v0;
            float3 ft[3];
            tri_terms(qs[u], rel(T.y - 1 - v0, u0), rel(T.z - 1 - v0, u0), Ac, Vc, ft);
            float* q = buf + 9 * i;
#pragma unroll
            for (int k = 0; k < 3; ++k) { q[3 * k] = ft[k].x; q[3 * k + 1] = ft[k].y; q[3 * k + 2] = ft[k].z; }
        }
    }
    __syncthreads();
    if (tid < nv) {
        if constexpr (PRE) {
            ti_pre.sum(buf, 9, tb, a.te, fx, fy, fz);
        } else {
            TriInc ti;
            ti.load(a.ts, a.te, v0 + tid);
            ti.sum(buf, 9, tb, a.te, fx, fy, fz);
        }
    }
    f = make_float3(fx, fy, fz);
    if (a.fault) {
        if (tid < nv && !isfinite(fx + fy + fz)) bad |= GC_FAULT_NONFINITE;
        if (bad) atomicOr(a.fault, bad);
    }
}

__global__ __launch_bounds__(GC_CELL_BS) void k_cell_membrane(GcParams P, const float* __restrict__ bbox,
                                                              CellArgs ca, DefArgs a)
{
    __shared__ double sx[GC_CELL_NV], sy[GC_CELL_NV], sz[GC_CELL_NV];
    __shared__ float buf[GC_CELL_NE * 12];
    __shared__ double sa[GC_CELL_BS], sv[GC_CELL_BS];
    if (ca.gate && !*ca.gate) return;   // block-uniform
    const int c = blockIdx.x, tid = threadIdx.x;
    const int4 r = ca.range[c];
    const int v0 = r.x, nv = r.y, e0 = r.z, ne = r.w;
    const int tb = ca.betri[c].x, te = ca.betri[c].y;
    const float* bb = bbox;
    if (a.reps) {
        const GcRep& R = a.reps[a.cell_rep[c]];
        P = R.P;
        bb = R.bbox;
    }
    const PerBox b = load_box(P, bb);
    // Everything a vertex's thread reads from global memory besides the
    // records of the phases is loaded here, ahead of its use: the
    // incidence lists for the two per-vertex sums and the inputs of the
    // force update and corrector (written by the preceding passes).
    EdgeInc ei;
    TriInc ti;
    int slot = -1;
    tfloat3 F0 = tfloat3{0.f, 0.f, 0.f}, vh = F0;
    float rw = 0.f;
    // unwrapped FP64 positions (k2_mgather's values) of the cell's vertices
    if (tid < nv) {
        const int g = v0 + tid;
        ei.load(a.es, a.ee, g);
        ti.load(a.ts, a.te, g);
        if (a.mpos) {
            sx[tid] = a.mpos[g]; sy[tid] = a.mpos[a.nm + g]; sz[tid] = a.mpos[2 * a.nm + g];
            if (a.force) slot = a.curpos[P.nc + g];
        } else {
            slot = a.curpos[P.nc + g];
            const tfloat3 x = a.coor[slot];
            const tint3 n = a.nthro[slot];
            sx[tid] = (double)x.x + (b.per[0] ? (double)n.x * (double)b.L[0] : 0.0);
            sy[tid] = (double)x.y + (b.per[1] ? (double)n.y * (double)b.L[1] : 0.0);
            sz[tid] = (double)x.z + (b.per[2] ? (double)n.z * (double)b.L[2] : 0.0);
        }
        if (a.force) {
            F0 = a.force[slot];
            if (a.vout) { vh = a.velhalf[slot]; rw = a.velrho[slot].w; }
        }
    }
    __syncthreads();
    float3 f;
    cell_force<true>(P, ca, a, c, tid, v0, nv, e0, ne, tb, te, true, sx, sy, sz, buf, sa, sv, f, ei, ti);
    if (tid >= nv) return;
    const int g = v0 + tid;
    a.defforce[g] = tfloat3{f.x, f.y, f.z};
    if (!a.force) return;   // membrane-only evaluation (substepping)
    tfloat3 F = F0;
    F.x += f.x; F.y += f.y; F.z += f.z;
    a.force[slot] = F;
    if (a.vout) {  // this is the vertex's last force term -> corrector
        const float h = 0.5f * P.dt * P.inv_m;
        a.vout[slot] = make_float4(vh.x + h * F.x, vh.y + h * F.y, vh.z + h * F.z, rw);
    }
}

// a pointer the compiler cannot see through (keeps loads inside a loop)
template <class T>
__device__ __forceinline__ T* opaque(T* p)
{
    asm volatile("" : "+s"(p));
    return p;
}

// k velocity-Verlet substeps of one cell's membrane under the fixed fluid
// force fe: per substep the kick and drift of k_ms_kick_drift, the membrane
// force of k_cell_membrane (geometry from the FP32-rounded positions) and
// the kick of k_ms_kick, with the same arithmetic, so the same bits.  The
// positions stay in LDS and each thread keeps its vertex's velocity and
// forces in registers between substeps.
__global__ __launch_bounds__(GC_CELL_BS) void k_cell_substeps(GcParams P, CellArgs ca, DefArgs a, SubArgs sa_)
{
    __shared__ double sx[GC_CELL_NV], sy[GC_CELL_NV], sz[GC_CELL_NV];
    __shared__ float buf[GC_CELL_NE * 12];
    __shared__ double sa[GC_CELL_BS], sv[GC_CELL_BS];
    if (ca.gate && !*ca.gate) return;   // block-uniform
    const int c = blockIdx.x, tid = threadIdx.x;
    const int4 r = ca.range[c];
    const int v0 = r.x, nv = r.y, e0 = r.z, ne = r.w;
    const int tb = ca.betri[c].x, te = ca.betri[c].y;
    const int nm = a.nm, g = v0 + tid;
    const bool own = tid < nv;
    float3 u = make_float3(0.f, 0.f, 0.f), fe = u, fm = u;
    if (own) {
        sx[tid] = sa_.mx_in[g]; sy[tid] = sa_.mx_in[nm + g]; sz[tid] = sa_.mx_in[2 * nm + g];
        const tfloat3 u0 = sa_.mv_in[g], a0 = sa_.fe[g], b0 = sa_.fm_in[g];
        u = make_float3(u0.x, u0.y, u0.z);
        fe = make_float3(a0.x, a0.y, a0.z);
        fm = make_float3(b0.x, b0.y, b0.z);
    }
    const float dts = sa_.dts;
    for (int q = 1; q <= sa_.k; ++q) {
        if (own) {
            ms_kick(u, fe, fm, dts, sa_.inv_m);
            sx[tid] = ms_drift(sx[tid], dts, u.x);
            sy[tid] = ms_drift(sy[tid], dts, u.y);
            sz[tid] = ms_drift(sz[tid], dts, u.z);
            double* keep = q == sa_.k - 2 ? sa_.mx6 : q == sa_.k - 1 ? sa_.mx7 : nullptr;
            if (keep) { keep[g] = sx[tid]; keep[nm + g] = sy[tid]; keep[2 * nm + g] = sz[tid]; }
        }
        __syncthreads();
        // the topology is re-read every substep (grouped loads, mostly cache
        // hits) rather than held in registers across the loop
        DefArgs al = a;
        al.tri = opaque(al.tri); al.edge = opaque(al.edge);
        al.es = opaque(al.es); al.ee = opaque(al.ee); al.ts = opaque(al.ts); al.te = opaque(al.te);
        al.lmax = opaque(al.lmax); al.p = opaque(al.p); al.kp = opaque(al.kp); al.mexp = opaque(al.mexp);
        al.kb = opaque(al.kb); al.th0 = opaque(al.th0); al.ka = opaque(al.ka); al.A0c = opaque(al.A0c);
        al.kd = opaque(al.kd); al.reftriarea = opaque(al.reftriarea); al.kv = opaque(al.kv);
        al.V0c = opaque(al.V0c);
        cell_force<false>(P, ca, al, c, tid, v0, nv, e0, ne, tb, te, q == sa_.k, sx, sy, sz, buf, sa, sv, fm,
                          EdgeInc{}, TriInc{});
        if (own) ms_kick(u, fe, fm, dts, sa_.inv_m);
        __syncthreads();
    }
    if (!own) return;
    const double x = sx[tid], y = sy[tid], z = sz[tid];
    sa_.mx[g] = x; sa_.mx[nm + g] = y; sa_.mx[2 * nm + g] = z;
    sa_.mcoor[g] = tfloat3{(float)x, (float)y, (float)z};
    sa_.mv[g] = tfloat3{u.x, u.y, u.z};
    sa_.fm[g] = tfloat3{fm.x, fm.y, fm.z};
    if (sa_.d2x) {   // this thread's own earlier stores
        const double* x6 = sa_.mx6;
        const double* x7 = sa_.mx7;
        const double x8[3] = {x, y, z};
        float c[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const size_t i = (size_t)d * nm + g;
            c[d] = (float)((x8[d] - x7[i]) - (x7[i] - x6[i]));
        }
        sa_.d2x[g] = tfloat3{c[0], c[1], c[2]};
    }
}

__global__ void k_sincos(int n, const float* __restrict__ th, float* __restrict__ c, float* __restrict__ s)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sincosf(th[i], &s[i], &c[i]);
}

bool cell_ranges(int nm, int nedge, int ntri, int ncell, const int* celltab, const tint2* betri,
                 const int6* edge, const tint3* tri, int4* range)
{
    if (ncell <= 0) return false;
    std::vector<int> vfirst(ncell, -1), vcount(ncell, 0), efirst(ncell, -1), ecount(ncell, 0);
    int prev = -1;
    for (int v = 0; v < nm; ++v) {
        const int c = celltab[v];
        if (c < 0 || c >= ncell || c < prev) return false;
        if (vfirst[c] < 0) vfirst[c] = v;
        vcount[c]++;
        prev = c;
    }
    prev = -1;
    for (int e = 0; e < nedge; ++e) {
        const int* q = &edge[e].x;
        if (q[0] < 1 || q[0] > nm) return false;
        const int c = celltab[q[0] - 1];
        if (c < prev) return false;
        for (int k = 1; k < 4; ++k)
            if (q[k] < 1 || q[k] > nm || celltab[q[k] - 1] != c) return false;
        if (efirst[c] < 0) efirst[c] = e;
        ecount[c]++;
        prev = c;
    }
    std::vector<char> covered(ntri, 0);
    for (int c = 0; c < ncell; ++c) {
        const int tb = betri[c].x, te = betri[c].y;
        if (vcount[c] < 1 || vcount[c] > GC_CELL_NV || ecount[c] > GC_CELL_NE) return false;
        if (tb < 0 || te < tb || te >= ntri || te - tb + 1 > GC_CELL_NT) return false;
        for (int t = tb; t <= te; ++t) {
            if (covered[t]) return false;
            covered[t] = 1;
        }
        for (int t = tb; t <= te; ++t) {
            const int* q = &tri[t].x;
            for (int k = 0; k < 3; ++k)
                if (q[k] < 1 || q[k] > nm || celltab[q[k] - 1] != c) return false;
        }
        range[c] = make_int4(vfirst[c], vcount[c], efirst[c] < 0 ? 0 : efirst[c], ecount[c]);
    }
    for (int t = 0; t < ntri; ++t)
        if (!covered[t]) return false;
    return true;
}

// ---- CalCellAggForce ------------------------------------------------------
__global__ __launch_bounds__(BS) void k_agg_force(int n, GcParams P, const tfloat3* __restrict__ coor,
    tfloat3* __restrict__ aggforce, const tint3* __restrict__ cc, const int* __restrict__ ncube,
    const int* __restrict__ sortpart, const int* __restrict__ celltab, const float* __restrict__ outer,
    const tint2* __restrict__ be, tfloat3* __restrict__ force)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int id = sortpart[p];
    if (id < P.nc) return;
    int v = id - P.nc;
    int cell = celltab[v];
    GcGridV g;
    g.nc[0] = ncube[0]; g.nc[1] = ncube[1]; g.nc[2] = ncube[2];
    g.L[0] = outer[0]; g.L[1] = outer[1]; g.L[2] = outer[2];
    tfloat3 xi = coor[p];
    tint3 c = cc[p];
    float fx = 0.f, fy = 0.f, fz = 0.f;
    const float tb = 2.0f * P.beta * P.De;
    visit_ranges(g, P.per, be, c.x, c.y, c.z, [&](int jb, int je, float sx, float sy, float sz) {
        for (int j = jb; j < je; ++j) {
            int idj = sortpart[j];
            if (idj < P.nc || celltab[idj - P.nc] == cell) continue;
            tfloat3 xj = coor[j];
            float rx = pdiff(xi.x, xj.x, sx), ry = pdiff(xi.y, xj.y, sy), rz = pdiff(xi.z, xj.z, sz);
            float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 >= P.mcut2 || r2 == 0.f) continue;
            float r = sqrtf(r2);
            float ex = expf(P.beta * (P.r0 - r));
            float fm = tb * (ex * ex - ex) / r;
            fx += fm * rx; fy += fm * ry; fz += fm * rz;
        }
    });
    aggforce[v] = tfloat3{fx, fy, fz};
    tfloat3 F = force[p];
    F.x += fx; F.y += fy; F.z += fz;
    force[p] = F;
}

}  // namespace

#define GC_PARAMS_OR_RETURN(P)          \
    GcParams P;                         \
    do { int _s = params(&P); if (_s) return _s; } while (0)

extern "C" int gc_InvalidateTopology(void)
{
    int st = hip_check(hipStreamSynchronize(stream()), "InvalidateTopology sync");
    free_topo();
    return st;
}

extern "C" int gc_CalGeoPara(long stepcount, int ntri, int ncell, float* d_bbox, int* d_curpos,
                             tint3* d_tri, tint2* d_beginendtri, tint3* d_nthrobound,
                             tfloat3* d_totalcoor, tfloat3* d_trinordir, float* d_triarea,
                             float* d_reftriarea, tfloat3* d_tricenter, float* d_refarea_para,
                             float* d_refvol_para, float* d_area_para, float* d_vol_para,
                             float* d_area, float* d_vol, float* d_pararea)
{
    if (ncell <= 0) return GC_OK;  // the reference calls it with nCell == 0 too
    GC_PARAMS_OR_RETURN(P);
    if (ntri <= 0) return set_error(GC_ERR_INVALID, "CalGeoPara: ntri <= 0 with cells");
    if (!d_bbox || !d_curpos || !d_tri || !d_beginendtri || !d_nthrobound || !d_totalcoor ||
        !d_trinordir || !d_triarea || !d_reftriarea || !d_tricenter || !d_refarea_para ||
        !d_refvol_para || !d_area_para || !d_vol_para || !d_area || !d_vol)
        return set_error(GC_ERR_INVALID, "CalGeoPara: null pointer");
    hipStream_t s = stream();
    k_geo_cell<<<ncell, GC_GEO_BS, 0, s>>>(P, stepcount, nullptr, nullptr, d_bbox, d_curpos, d_tri, d_beginendtri,
                                    d_nthrobound, d_totalcoor, d_trinordir, d_triarea,
                                    d_reftriarea, d_tricenter, d_refarea_para, d_refvol_para,
                                    d_area_para, d_vol_para, d_area, d_vol);
    GC_LAUNCH_CHECK("CalGeoPara");
    if (d_pararea) {
        bool cached = g_topo.tri_ptr == (const void*)d_tri && g_topo.ntri == ntri && g_topo.nm == P.nm;
        int st = cached ? GC_OK : ensure_topology(P.nm, 0, nullptr, ntri, d_tri);
        if (st) return st;
        k_par_area<<<blocks_for(P.nm, BS), BS, 0, s>>>(P.nm, g_topo.d_ts, g_topo.d_te, d_triarea,
                                                       d_pararea);
        GC_LAUNCH_CHECK("CalGeoPara pararea");
    }
    return GC_OK;
}

extern "C" int gc_SetMechPara(int ncell, int nedge, int6* d_edge, int* d_curpos,
                              float* d_refedgelen, float* d_maxrefedgelen, float* d_bendmodpara,
                              float* d_shearmodpara, float* d_hbending, float* d_powerindex,
                              float* d_perlenWLC, float* d_hpow, float* d_reftriangle,
                              tfloat3* d_trinordir, tfloat3* d_tricente, tfloat3* d_coor)
{
    (void)d_trinordir; (void)d_tricente;
    if (ncell <= 0) return GC_OK;
    GC_PARAMS_OR_RETURN(P);
    if (nedge <= 0) return set_error(GC_ERR_INVALID, "SetMechPara: nedge <= 0 with cells");
    if (!d_edge || !d_curpos || !d_refedgelen || !d_maxrefedgelen || !d_bendmodpara ||
        !d_shearmodpara || !d_hbending || !d_powerindex || !d_perlenWLC || !d_hpow ||
        !d_reftriangle || !d_coor)
        return set_error(GC_ERR_INVALID, "SetMechPara: null pointer");
    k_mech_para<<<blocks_for(nedge, BS), BS, 0, stream()>>>(P, nedge, d_edge, d_curpos, d_coor,
        d_refedgelen, d_maxrefedgelen, d_bendmodpara, d_shearmodpara, d_hbending, d_powerindex,
        d_perlenWLC, d_hpow, d_reftriangle);
    GC_LAUNCH_CHECK("SetMechPara");
    return GC_OK;
}

extern "C" int gc_CalCellDefForce(int ncell, unsigned nedge, unsigned ntri, unsigned ntotalpar,
                                  unsigned ncellpar, float* d_bbox, int6* d_edge, tint3* d_tri,
                                  tfloat3* d_coor, tfloat3* h_defforce, tfloat3* d_defforce,
                                  tint3* d_nthrobound, float* d_maxrefedgelen, float* d_powerindex,
                                  float* d_perlenWLC, float* d_hpow, tfloat3* d_trinordir,
                                  tfloat3* d_tricenter, float* d_reftriangle, float* d_hbending,
                                  float* d_triarea, float* d_reftriarea, float* d_hlarea_para,
                                  float* d_hgarea_para, float* d_area_para, float* d_refarea_para,
                                  float* d_hvol_para, float* d_vol_para, float* d_refvol_para,
                                  tfloat3* d_totalforce, int* d_sortpart, int* d_curpos)
{
    (void)ntotalpar; (void)d_trinordir; (void)d_tricenter; (void)d_triarea; (void)d_sortpart;
    if (ncell <= 0) return GC_OK;
    GC_PARAMS_OR_RETURN(P);
    if ((int)ncellpar != P.nm) return set_error(GC_ERR_INVALID, "CalCellDefForce: ncellpar != nCellPar");
    if (!d_bbox || !d_edge || !d_tri || !d_coor || !d_defforce || !d_nthrobound ||
        !d_maxrefedgelen || !d_powerindex || !d_perlenWLC || !d_hpow || !d_reftriangle ||
        !d_hbending || !d_reftriarea || !d_hlarea_para || !d_hgarea_para || !d_area_para ||
        !d_refarea_para || !d_hvol_para || !d_vol_para || !d_refvol_para || !d_totalforce ||
        !d_curpos)
        return set_error(GC_ERR_INVALID, "CalCellDefForce: null pointer");
    int st = ensure_topology(P.nm, (int)nedge, d_edge, (int)ntri, d_tri);
    if (st) return st;
    DefArgs a;
    a.edge = d_edge; a.tri = d_tri; a.coor = d_coor; a.nthro = d_nthrobound; a.curpos = d_curpos;
    a.lmax = d_maxrefedgelen; a.mexp = d_powerindex; a.p = d_perlenWLC; a.kp = d_hpow;
    a.th0 = d_reftriangle; a.kb = d_hbending; a.reftriarea = d_reftriarea; a.kd = d_hlarea_para;
    a.ka = d_hgarea_para; a.Ac = d_area_para; a.A0c = d_refarea_para; a.kv = d_hvol_para;
    a.Vc = d_vol_para; a.V0c = d_refvol_para;
    a.es = g_topo.d_es; a.ee = g_topo.d_ee; a.ts = g_topo.d_ts; a.te = g_topo.d_te;
    a.defforce = d_defforce; a.force = d_totalforce;
    k_def_force<<<xcd_grid(blocks_for(P.nm, BS)), BS, 0, stream()>>>(P, P.nm, d_bbox, a);
    GC_LAUNCH_CHECK("CalCellDefForce");
    if (h_defforce) {
        st = hip_check(hipMemcpyAsync(h_defforce, d_defforce, sizeof(tfloat3) * (size_t)P.nm,
                                      hipMemcpyDeviceToHost, stream()), "DefForce D2H");
        if (!st) st = hip_check(hipStreamSynchronize(stream()), "DefForce sync");
    }
    return st;
}

extern "C" int gc_CalCellAggForce(unsigned ncell, unsigned n, unsigned ncellpar,
                                  tfloat3* d_totalcoor, tfloat3* d_aggforce, tint3* d_cc,
                                  int* d_ncube, int* d_sortpart, int* d_celltab,
                                  float* d_outerreigon, tint2* d_beginendcube,
                                  tfloat3* d_totalforce, float* d_pararea)
{
    (void)d_pararea;
    if (ncell == 0) return GC_OK;
    GC_PARAMS_OR_RETURN(P);
    if ((int)ncellpar != P.nm) return set_error(GC_ERR_INVALID, "CalCellAggForce: ncellpar != nCellPar");
    if (!d_totalcoor || !d_aggforce || !d_cc || !d_ncube || !d_sortpart || !d_celltab ||
        !d_outerreigon || !d_beginendcube || !d_totalforce)
        return set_error(GC_ERR_INVALID, "CalCellAggForce: null pointer");
    if (n == 0) return GC_OK;
    k_agg_force<<<blocks_for(n, BS), BS, 0, stream()>>>((int)n, P, d_totalcoor, d_aggforce, d_cc,
        d_ncube, d_sortpart, d_celltab, d_outerreigon, d_beginendcube, d_totalforce);
    GC_LAUNCH_CHECK("CalCellAggForce");
    return GC_OK;
}
