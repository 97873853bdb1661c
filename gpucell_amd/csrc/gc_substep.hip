// gc_substep.hip -- membrane time stepping of the engine: plain velocity-
// Verlet substepping of the stiff membrane inside each fluid step, and the
// learned substepper that predicts the substeps with the MLP (gc_mlp.hip),
// checks the prediction against the physics and falls back to plain
// substepping.  No counterpart in the reference, whose membrane vertices
// move with the fluid step (Particle::UpdatePar, src/Particle.cu:1310-1329).
//
// The membrane runs as its own small system: unwrapped FP64 positions in
// id order (x[nm], y[nm], z[nm]), FP32 velocities, the fluid force of the
// step held fixed, and the membrane kernels of gc_membrane.hip evaluated on
// a parameter block with no fluid (nc = 0) and no periodic axes, so a
// vertex's slot is its index and its positions are read unwrapped.
//
// Per engine step (after the fluid forces):
//   fext = force of the vertex's slot (SDPD + Morse + body force)
//   plain, k substeps of dts = dt/k:
//     v += dts/2 (fext + fmem)/m;  x += dts v;  fmem = F(x);  v += dts/2 (fext + fmem)/m
//   surrogate: features -> MLP -> x1..x8; fmem = F(x8); v8 = (x8 - x7)/dts +
//     dts/2 a8; residual R; R > threshold -> restore and run plain.
//   write back: wrapped positions, image counters, velocities into the
//   engine's slot arrays.
#include "gc_engine_state.h"
#include "../../include/gpucell/surrogate.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

using namespace gc;

namespace {

constexpr int BS = GC_BS;
constexpr int KSUB = 8;          // substeps the network predicts
constexpr int NF = GC_MLP_IN, NO = GC_MLP_OUT;

__device__ inline float3 ld3(const tfloat3& a) { return make_float3(a.x, a.y, a.z); }
__device__ inline float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float3 cross3(float3 a, float3 b)
{
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline float3 scl3(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ inline float norm3(float3 a) { return sqrtf(dot3(a, a)); }

// engine slots -> unwrapped FP64 positions and velocities in id order
__global__ void k_ms_init(int nm, int nc, int N, const int* __restrict__ curpos,
                          const double* __restrict__ pos64, const tint3* __restrict__ nt,
                          const float4* __restrict__ vout, const float* __restrict__ bbox, int3 per,
                          double* __restrict__ mx, tfloat3* __restrict__ mv, tfloat3* __restrict__ mcoor)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nm) return;
    int s = curpos[nc + v];
    tint3 n = nt[s];
    double L0 = (double)(bbox[1] - bbox[0]), L1 = (double)(bbox[3] - bbox[2]), L2 = (double)(bbox[5] - bbox[4]);
    double x = pos64[s] + (per.x ? n.x * L0 : 0.0);
    double y = pos64[N + s] + (per.y ? n.y * L1 : 0.0);
    double z = pos64[2 * N + s] + (per.z ? n.z * L2 : 0.0);
    mx[v] = x; mx[nm + v] = y; mx[2 * nm + v] = z;
    mcoor[v] = tfloat3{(float)x, (float)y, (float)z};
    float4 q = vout[s];
    mv[v] = tfloat3{q.x, q.y, q.z};
}

// the fluid force of each membrane vertex (slot order -> id order); with
// mx0 also a copy of the start state (positions, velocity, membrane force)
__global__ void k_ms_fext(int nm, int nc, const int* __restrict__ curpos, const tfloat3* __restrict__ force,
                          tfloat3* __restrict__ fe, const double* __restrict__ mx, const tfloat3* __restrict__ mv,
                          const tfloat3* __restrict__ fm, double* __restrict__ mx0, tfloat3* __restrict__ mv0,
                          tfloat3* __restrict__ fm0)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nm) return;
    fe[v] = force[curpos[nc + v]];
    if (!mx0) return;
    mx0[v] = mx[v]; mx0[nm + v] = mx[nm + v]; mx0[2 * nm + v] = mx[2 * nm + v];
    mv0[v] = mv[v];
    fm0[v] = fm[v];
}

__global__ void k_ms_kick_drift(int nm, float dts, float inv_m, double* __restrict__ mx, tfloat3* __restrict__ mv,
                                const tfloat3* __restrict__ fe, const tfloat3* __restrict__ fm,
                                tfloat3* __restrict__ mcoor, const int* __restrict__ gate)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nm || (gate && !*gate)) return;
    float3 u = ld3(mv[v]);
    ms_kick(u, ld3(fe[v]), ld3(fm[v]), dts, inv_m);
    mv[v] = tfloat3{u.x, u.y, u.z};
    const double x = ms_drift(mx[v], dts, u.x), y = ms_drift(mx[nm + v], dts, u.y), z = ms_drift(mx[2 * nm + v], dts, u.z);
    mx[v] = x; mx[nm + v] = y; mx[2 * nm + v] = z;
    mcoor[v] = tfloat3{(float)x, (float)y, (float)z};
}

__global__ void k_ms_kick(int nm, float dts, float inv_m, tfloat3* __restrict__ mv, const tfloat3* __restrict__ fe,
                          const tfloat3* __restrict__ fm, const int* __restrict__ gate)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nm || (gate && !*gate)) return;
    float3 u = ld3(mv[v]);
    ms_kick(u, ld3(fe[v]), ld3(fm[v]), dts, inv_m);
    mv[v] = tfloat3{u.x, u.y, u.z};
}

// membrane state -> engine slots: wrapped FP64 position, image counters,
// FP32 coordinates, velocity, membrane force
// skin mode (woff != NULL): the wraps since the last list rebuild join the
// vertex's wrap code, and the statistics of its displacement D since that
// rebuild (in the rebuild's image frame, from rbpos) the group's maxima
// bmax[DSTAT_N blockIdx.x ..]: |D|^2, then per axis max(D, 0), max(-D, 0)
__global__ __launch_bounds__(BS) void k_ms_end(int nm, int nc, int N, const int* __restrict__ curpos,
                         const double* __restrict__ mx,
                         const tfloat3* __restrict__ mv, const tfloat3* __restrict__ fm,
                         const float* __restrict__ bbox, int3 per, double* __restrict__ pos64,
                         tint3* __restrict__ nt, tfloat3* __restrict__ coor, float4* __restrict__ vout,
                         tfloat3* __restrict__ defforce, int* __restrict__ flag,
                         unsigned char* __restrict__ woff, const tfloat3* __restrict__ rbpos,
                         float* __restrict__ bmax)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v == 0) flag[2] += 1;   // steps in this mode
    const bool live = v < nm;
    float st[DSTAT_N] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (live) {
    int s = curpos[nc + v];
    double u[3] = {mx[v], mx[nm + v], mx[2 * nm + v]};
    int img[3] = {0, 0, 0};
    const int pp[3] = {per.x, per.y, per.z};
    const tint3 n0 = nt[s];
    const int im0[3] = {n0.x, n0.y, n0.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!pp[d]) continue;
        double lo = bbox[2 * d], L = (double)(bbox[2 * d + 1] - bbox[2 * d]);
        double k = floor((u[d] - lo) / L);
        double w = u[d] - k * L;
        if (w >= (double)bbox[2 * d + 1]) { w -= L; k += 1.0; }
        if (w < lo) { w += L; k -= 1.0; }
        u[d] = w;
        img[d] = (int)k;
    }
    if (woff) {
        int c = woff[s];
        if (img[0] != im0[0] || img[1] != im0[1] || img[2] != im0[2]) {
            // wraps since the last list rebuild (image code, per axis -1..1)
            int k3[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int b = (c >> (2 * d)) & 3;
                int k = ((b & 2) ? 1 : -(b & 1)) + (img[d] - im0[d]);
                k3[d] = k < -1 ? -1 : (k > 1 ? 1 : k);
            }
            auto enc = [](int k) { return k < 0 ? 1 : (k > 0 ? 2 : 0); };
            c = enc(k3[0]) | (enc(k3[1]) << 2) | (enc(k3[2]) << 4);
            woff[s] = (unsigned char)c;
        }
        const tfloat3 r0 = rbpos[s];
        const float r0v[3] = {r0.x, r0.y, r0.z};
        float d2 = 0.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int b = (c >> (2 * d)) & 3;
            const float k = (b & 2) ? 1.0f : ((b & 1) ? -1.0f : 0.0f);
            const float dx = ((float)u[d] + k * (bbox[2 * d + 1] - bbox[2 * d])) - r0v[d];
            d2 += dx * dx;
            st[1 + 2 * d] = fmaxf(dx, 0.0f);
            st[2 + 2 * d] = fmaxf(-dx, 0.0f);
        }
        st[0] = d2 <= 3.0e38f ? d2 : 3.0e38f;
        if (!(st[0] < 3.0e38f))
            for (int k = 1; k < DSTAT_N; ++k) st[k] = 3.0e38f;
    }
    pos64[s] = u[0]; pos64[N + s] = u[1]; pos64[2 * N + s] = u[2];
    nt[s] = tint3{img[0], img[1], img[2]};
    coor[s] = tfloat3{(float)u[0], (float)u[1], (float)u[2]};
    tfloat3 q = mv[v];
    float4 o = vout[s];
    vout[s] = make_float4(q.x, q.y, q.z, o.w);
    defforce[v] = fm[v];
    }
    if (woff) {   // the group's displacement statistics
        __shared__ float sm[BS / 64][DSTAT_N];
#pragma unroll
        for (int k = 0; k < DSTAT_N; ++k) {
            float m = st[k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
            if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6][k] = m;
        }
        __syncthreads();
        if (threadIdx.x < DSTAT_N) {
            float b = 0.0f;
            for (int q = 0; q < BS / 64; ++q) b = fmaxf(b, sm[q][threadIdx.x]);
            bmax[DSTAT_N * blockIdx.x + threadIdx.x] = b;
        }
    }
}

// leaving substepping: the slot force becomes the total force again
__global__ void k_ms_addforce(int nm, int nc, const int* __restrict__ curpos, const tfloat3* __restrict__ fm,
                              tfloat3* __restrict__ force)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nm) return;
    int s = curpos[nc + v];
    tfloat3 f = force[s], m = fm[v];
    force[s] = tfloat3{f.x + m.x, f.y + m.y, f.z + m.z};
}

struct FeatArgs {
    const double* mx; const tfloat3* mv; const tfloat3* fe; const tfloat3* fm;
    const int* ring; const float* lref;
    const int* ts; const int* te; const float* triarea; const float* reftriarea;
    const float* Ac; const float* A0c; const float* Vc; const float* V0c;
    int nm; float dt; float inv_m;
    const float* norm;   // NULL or mean[36] std[36] ...
    float* frame; float* feat;
};

__device__ inline float3 mrel(const double* mx, int nm, int j, double ox, double oy, double oz)
{
    return make_float3((float)(mx[j] - ox), (float)(mx[nm + j] - oy), (float)(mx[2 * nm + j] - oz));
}

// rows of w floats formed per thread in st[thread (w + 1) + c] -> out as
// contiguous runs (the rows of consecutive vertices are adjacent)
__device__ __forceinline__ void store_rows(const float* st, int w, float* __restrict__ out, int rows)
{
    __syncthreads();
    for (int i = threadIdx.x; i < rows * w; i += BS) out[i] = st[(i / w) * (w + 1) + i % w];
    __syncthreads();
}

// rows of W floats from in (contiguous) -> st[row (W + 1) + c]; the caller
// synchronises before reading st.  All W loads of a thread are issued before
// the first store (clamped addresses, no branch around a load).
template <int W>
__device__ __forceinline__ void load_rows(const float* __restrict__ in, float* st, int rows)
{
    const int n = rows * W;
    float t[W];
#pragma unroll
    for (int k = 0; k < W; ++k) t[k] = in[min((int)threadIdx.x + k * BS, n - 1)];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int i = threadIdx.x + k * BS;
        if (i < n) st[(i / W) * (W + 1) + i % W] = t[k];
    }
}

// 36 features of vertex v in its frame (e1, e2, n):
//  0-17 1-ring edge vectors / l (CCW from the neighbour that defines e1;
//       zeros for a missing sixth neighbour)
//  18-20 v dt / l; 21-23 fext dt^2/(m l); 24-26 fmem dt^2/(m l);
//  27-29 (mean ring velocity - v) dt / l;
//  30 A/A0 - 1 (cell), 31 V/V0 - 1 (cell), 32 local area / rest - 1,
//  33 1 for a five-neighbour vertex, 34 mean ring edge / l - 1, 35 zero.
// Frame rows (e1, e2, n) go to frame[9 v].
__global__ __launch_bounds__(BS) void k_ms_features(FeatArgs a)
{
    __shared__ float st[BS * (NF + 1)];
    const int nm = a.nm;
    const int v0 = blockIdx.x * BS, v = v0 + (int)threadIdx.x;
    const int rows = min(BS, nm - v0);
    float f[NF], fr[9];
    if (v < nm) {
        const double ox = a.mx[v], oy = a.mx[nm + v], oz = a.mx[2 * nm + v];
        int r[6];
        int deg = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) { r[k] = a.ring[6 * v + k]; deg += r[k] >= 0; }
        float3 d[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) d[k] = r[k] >= 0 ? mrel(a.mx, nm, r[k], ox, oy, oz) : make_float3(0.f, 0.f, 0.f);
        float3 n = make_float3(0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            if (k >= deg) break;
            float3 c = cross3(d[k], d[(k + 1) % deg]);
            n.x += c.x; n.y += c.y; n.z += c.z;
        }
        n = scl3(n, 1.0f / fmaxf(norm3(n), 1e-30f));
        float3 e1 = sub3(d[0], scl3(n, dot3(d[0], n)));
        e1 = scl3(e1, 1.0f / fmaxf(norm3(e1), 1e-30f));
        float3 e2 = cross3(n, e1);
        fr[0] = e1.x; fr[1] = e1.y; fr[2] = e1.z; fr[3] = e2.x; fr[4] = e2.y; fr[5] = e2.z;
        fr[6] = n.x; fr[7] = n.y; fr[8] = n.z;
        const float il = 1.0f / a.lref[v];
        float mean_len = 0.f;
        float3 vr = make_float3(0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            f[3 * k] = dot3(d[k], e1) * il; f[3 * k + 1] = dot3(d[k], e2) * il; f[3 * k + 2] = dot3(d[k], n) * il;
            if (r[k] >= 0) {
                mean_len += norm3(d[k]);
                float3 w = ld3(a.mv[r[k]]);
                vr.x += w.x; vr.y += w.y; vr.z += w.z;
            }
        }
        const float3 vv = ld3(a.mv[v]);
        vr = scl3(vr, 1.0f / (float)deg);
        vr = sub3(vr, vv);
        const float sv = a.dt * il, sf = a.dt * a.dt * a.inv_m * il;
        const float3 fe = ld3(a.fe[v]), fm = ld3(a.fm[v]);
        f[18] = dot3(vv, e1) * sv; f[19] = dot3(vv, e2) * sv; f[20] = dot3(vv, n) * sv;
        f[21] = dot3(fe, e1) * sf; f[22] = dot3(fe, e2) * sf; f[23] = dot3(fe, n) * sf;
        f[24] = dot3(fm, e1) * sf; f[25] = dot3(fm, e2) * sf; f[26] = dot3(fm, n) * sf;
        f[27] = dot3(vr, e1) * sv; f[28] = dot3(vr, e2) * sv; f[29] = dot3(vr, n) * sv;
        const int t0 = a.te[a.ts[v]] >> 2;
        f[30] = a.Ac[t0] / a.A0c[t0] - 1.0f;
        f[31] = a.Vc[t0] / a.V0c[t0] - 1.0f;
        float la = 0.f, lr = 0.f;
        for (int q = a.ts[v]; q < a.ts[v + 1]; ++q) {
            int t = a.te[q] >> 2;
            la += a.triarea[t];
            lr += a.reftriarea[t];
        }
        f[32] = la / lr - 1.0f;
        f[33] = deg == 5 ? 1.0f : 0.0f;
        f[34] = mean_len / (float)deg * il - 1.0f;
        f[35] = 0.0f;
        float* my = st + 10 * threadIdx.x;
#pragma unroll
        for (int c = 0; c < 9; ++c) my[c] = fr[c];
    }
    store_rows(st, 9, a.frame + 9 * (size_t)v0, rows);
    if (v < nm) {
        float* my = st + (NF + 1) * threadIdx.x;
#pragma unroll
        for (int c = 0; c < NF; ++c) my[c] = a.norm ? (f[c] - a.norm[c]) / a.norm[NF + c] : f[c];
    }
    store_rows(st, NF, a.feat + (size_t)NF * v0, rows);
}

// recording: target block of substep s (1-based) of every vertex
__global__ void k_ms_target(int nm, int s, float dts, const double* __restrict__ mx, const double* __restrict__ mx0,
                            const tfloat3* __restrict__ mv0, const float* __restrict__ frame,
                            const float* __restrict__ lref, float* __restrict__ y)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nm) return;
    tfloat3 u = mv0[v];
    double t = (double)s * dts;
    float dx = (float)(mx[v] - mx0[v] - t * u.x);
    float dy = (float)(mx[nm + v] - mx0[nm + v] - t * u.y);
    float dz = (float)(mx[2 * nm + v] - mx0[2 * nm + v] - t * u.z);
    const float* fr = frame + 9 * v;
    const float il = 1.0f / lref[v];
    float* o = y + (size_t)NO * v + 3 * (s - 1);
    o[0] = (fr[0] * dx + fr[1] * dy + fr[2] * dz) * il;
    o[1] = (fr[3] * dx + fr[4] * dy + fr[5] * dz) * il;
    o[2] = (fr[6] * dx + fr[7] * dy + fr[8] * dz) * il;
}

// surrogate: predicted substep displacements -> x8 (into mx), d2x, the
// half-step velocity (x8 - x7)/dts into mv
__global__ __launch_bounds__(BS) void k_ms_apply(int nm, float dts, const float* __restrict__ pred,
                           const float* __restrict__ norm,
                           const float* __restrict__ frame, const float* __restrict__ lref,
                           const double* __restrict__ mx0, const tfloat3* __restrict__ mv0,
                           double* __restrict__ mx, tfloat3* __restrict__ mv, tfloat3* __restrict__ mcoor,
                           tfloat3* __restrict__ d2x)
{
    // the block's prediction and frame rows, read as contiguous runs
    __shared__ float sp[BS * (NO + 1)], sf[BS * 10];
    const int v0 = blockIdx.x * BS, v = v0 + (int)threadIdx.x;
    const int rows = min(BS, nm - v0);
    load_rows<NO>(pred + (size_t)NO * v0, sp, rows);
    load_rows<9>(frame + 9 * (size_t)v0, sf, rows);
    __syncthreads();
    if (v >= nm) return;
    const float* fr = sf + 10 * threadIdx.x;
    const float* pv = sp + (NO + 1) * threadIdx.x;
    const float L = lref[v];
    float3 del[3];   // displacements of substeps 6, 7, 8 (minus the ballistic part)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int s = 5 + q;   // 0-based substep index 5, 6, 7
        float y[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            int o = 3 * s + c;
            float p = pv[o];
            y[c] = norm ? p * norm[2 * NF + NO + o] + norm[2 * NF + o] : p;
        }
        del[q] = make_float3(L * (y[0] * fr[0] + y[1] * fr[3] + y[2] * fr[6]),
                             L * (y[0] * fr[1] + y[1] * fr[4] + y[2] * fr[7]),
                             L * (y[0] * fr[2] + y[1] * fr[5] + y[2] * fr[8]));
    }
    tfloat3 u = mv0[v];
    const double t8 = (double)KSUB * dts;
    double x = mx0[v] + t8 * u.x + del[2].x, y = mx0[nm + v] + t8 * u.y + del[2].y,
           z = mx0[2 * nm + v] + t8 * u.z + del[2].z;
    mx[v] = x; mx[nm + v] = y; mx[2 * nm + v] = z;
    mcoor[v] = tfloat3{(float)x, (float)y, (float)z};
    const float idt = 1.0f / dts;
    mv[v] = tfloat3{u.x + (del[2].x - del[1].x) * idt, u.y + (del[2].y - del[1].y) * idt,
                    u.z + (del[2].z - del[1].z) * idt};
    d2x[v] = tfloat3{del[2].x - 2.0f * del[1].x + del[0].x, del[2].y - 2.0f * del[1].y + del[0].y,
                     del[2].z - 2.0f * del[1].z + del[0].z};
}

// plain trajectory: d2x = x8 - 2 x7 + x6 from the substep positions
__global__ void k_ms_d2x(int nm, const double* __restrict__ x8, const double* __restrict__ x7,
                         const double* __restrict__ x6, tfloat3* __restrict__ d2x)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nm) return;
    float c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        size_t i = (size_t)d * nm + v;
        c[d] = (float)((x8[i] - x7[i]) - (x7[i] - x6[i]));
    }
    d2x[v] = tfloat3{c[0], c[1], c[2]};
}

// residual partial sums (block tree, fixed order); with kick, first the
// surrogate's end kick v8 = v7.5 + dts/2 a8
__global__ __launch_bounds__(BS) void k_ms_finish(int nm, int kick, float dts, float inv_m, tfloat3* __restrict__ mv,
    const tfloat3* __restrict__ fe, const tfloat3* __restrict__ fm, const tfloat3* __restrict__ d2x,
    const float* __restrict__ lref, float* __restrict__ part)
{
    __shared__ float sh[BS];
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    float q = 0.f;
    if (v < nm) {
        tfloat3 a = fe[v], b = fm[v], d = d2x[v];
        float3 acc = make_float3((a.x + b.x) * inv_m, (a.y + b.y) * inv_m, (a.z + b.z) * inv_m);
        if (kick) {
            const float h = 0.5f * dts;
            tfloat3 u = mv[v];
            mv[v] = tfloat3{u.x + h * acc.x, u.y + h * acc.y, u.z + h * acc.z};
        }
        float3 e = scl3(acc, dts * dts);
        float3 dd = make_float3(d.x, d.y, d.z);
        float r = norm3(sub3(dd, e)) / lref[v];
        q = r * r;
    }
    sh[threadIdx.x] = q;
    __syncthreads();
    for (int o = BS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// out[0] = rms residual; with flag: flag[0] = fallback needed (R > thr or
// not a number), flag[1] counts the fallbacks.  One block: strided sums,
// then a tree (fixed order).  (A last-block reduction inside k_ms_finish
// needs a device-scope release per block, which writes back the L2: 10M
// 17 -> 250 us.)
constexpr int RES_BS = 1024;
__global__ __launch_bounds__(RES_BS) void k_ms_res_final(int nb, int nm, const float* __restrict__ part,
                                                         float* __restrict__ out, float thr, int* __restrict__ flag)
{
    __shared__ float sh[RES_BS];
    const int tid = threadIdx.x;
    float q = 0.f;
    for (int i = tid; i < nb; i += RES_BS) q += part[i];
    sh[tid] = q;
    __syncthreads();
    for (int o = RES_BS / 2; o > 0; o >>= 1) {
        if (tid < o) sh[tid] += sh[tid + o];
        __syncthreads();
    }
    if (tid != 0) return;
    const float r = sqrtf(sh[0] / (float)nm);
    out[0] = r;
    if (flag) {
        const int fb = !(r <= thr);
        flag[0] = fb;
        flag[1] += fb;
    }
}

// copies the substep state (positions, velocity, membrane force): the
// restore of the saved start state; with a gate only when *gate != 0
__global__ void k_ms_copy_state(int nm, const double* __restrict__ mx, const tfloat3* __restrict__ mv,
                                const tfloat3* __restrict__ fm, double* __restrict__ mx_to, tfloat3* __restrict__ mv_to,
                                tfloat3* __restrict__ fm_to, const int* __restrict__ gate)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nm || (gate && !*gate)) return;
    mx_to[v] = mx[v]; mx_to[nm + v] = mx[nm + v]; mx_to[2 * nm + v] = mx[2 * nm + v];
    mv_to[v] = mv[v];
    fm_to[v] = fm[v];
}

// device copy (graph-safe)
__global__ void k_ms_copy_d(int n, const double* __restrict__ a, double* __restrict__ b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    b[i] = a[i];
}

}  // namespace

namespace gc {

struct MembraneSub {
    int mode = 0, k = KSUB;
    gc_mlp* mlp = nullptr;
    float thr = 0.f;
    GcParams Pm;
    int nm = 0;
    double *mx = nullptr, *mx0 = nullptr, *mx6 = nullptr, *mx7 = nullptr;
    tfloat3 *mv, *mv0, *fm, *fm0, *fe, *mcoor, *d2x;
    int* mid; tint3* mnt;
    int* ring; float* lref;
    float *frame, *feat, *pred, *norm = nullptr, *part, *res;
    int* flag = nullptr;   // [0] fallback this step, [1] fallbacks, [2] steps (device counters)
    bool have_norm = false;
    float* rec_x = nullptr; float* rec_y = nullptr;
    long rec_cap = 0, rec_count = 0;
    long fallbacks = 0;   // decided on the host (three-pass membrane path)
    std::vector<void*> allocs;

    template <class T> int alloc(T** p, size_t n)
    {
        int st = dalloc(p, n);
        if (!st) allocs.push_back((void*)*p);
        return st;
    }
};

void ms_destroy(MembraneSub* m)
{
    if (!m) return;
    for (void* p : m->allocs) (void)hipFree(p);
    delete m;
}

}  // namespace gc

namespace {

int3 per3(const gc_engine* e) { return make_int3(e->ext.periodic[0], e->ext.periodic[1], e->ext.periodic[2]); }

// the membrane system's arguments of the membrane force kernels
DefArgs ms_defargs(gc_engine* e, MembraneSub* m)
{
    DefArgs a;
    a.edge = e->d_edge; a.tri = e->d_tri; a.coor = m->mcoor; a.nthro = m->mnt; a.curpos = m->mid;
    a.lmax = e->d_lmax; a.mexp = e->d_mexp; a.p = e->d_p; a.kp = e->d_kp; a.th0 = e->d_th0;
    a.cth0 = e->d_cth0; a.sth0 = e->d_sth0;
    a.kb = e->d_kb; a.reftriarea = e->d_reftriarea; a.kd = e->d_kd; a.ka = e->d_ka;
    a.Ac = e->d_area_para; a.A0c = e->d_refarea_para; a.kv = e->d_kv; a.Vc = e->d_vol_para;
    a.V0c = e->d_refvol_para; a.es = e->d_es; a.ee = e->d_ee; a.ts = e->d_ts; a.te = e->d_te;
    a.defforce = m->fm; a.force = nullptr;
    a.mpos = m->mx; a.nm = m->nm;
    a.fault = e->d_fault;
    return a;
}

// membrane force (and geometry) at the membrane system's positions; watch =
// false for the surrogate's predicted state (a bad prediction is rejected by
// its residual and replaced by plain substeps, not reported as a fault)
int ms_force(gc_engine* e, MembraneSub* m, const int* gate = nullptr, bool watch = true)
{
    hipStream_t s = e->s;
    DefArgs a = ms_defargs(e, m);
    if (!watch) a.fault = nullptr;
    if (use_cell_kernel(e)) {   // geometry from the FP32 positions, as k_geo_cell below
        CellArgs ca = cell_args(e, e->step, 1);
        ca.gate = gate;
        k_cell_membrane<<<e->ncell, GC_CELL_BS, 0, s>>>(m->Pm, e->d_bbox, ca, a);
        GC_LAUNCH_CHECK("membrane substep (per cell)");
        return GC_OK;
    }
    k_geo_cell<<<e->ncell, GC_GEO_BS, 0, s>>>(m->Pm, e->step, nullptr, nullptr, e->d_bbox, m->mid, e->d_tri,
        e->d_betri, m->mnt, m->mcoor, e->d_trinor, e->d_triarea, e->d_reftriarea, e->d_tricen,
        e->d_refarea_para, e->d_refvol_para, e->d_area_para, e->d_vol_para, e->d_area, e->d_vol);
    return membrane_force(e, m->Pm, a, s);
}

FeatArgs feat_args(gc_engine* e, MembraneSub* m, bool normalise)
{
    FeatArgs a;
    a.mx = m->mx; a.mv = m->mv; a.fe = m->fe; a.fm = m->fm; a.ring = m->ring; a.lref = m->lref;
    a.ts = e->d_ts; a.te = e->d_te; a.triarea = e->d_triarea; a.reftriarea = e->d_reftriarea;
    a.Ac = e->d_area_para; a.A0c = e->d_refarea_para; a.Vc = e->d_vol_para; a.V0c = e->d_refvol_para;
    a.nm = m->nm; a.dt = e->hreps[0].P.dt; a.inv_m = e->hreps[0].P.inv_m;
    a.norm = normalise && m->have_norm ? m->norm : nullptr;
    a.frame = m->frame; a.feat = m->feat;
    return a;
}

// back to the saved start state (only when *gate != 0 if gate is given)
int restore_start(gc_engine* e, MembraneSub* m, const int* gate = nullptr)
{
    k_ms_copy_state<<<blocks_for(m->nm, BS), BS, 0, e->s>>>(m->nm, m->mx0, m->mv0, m->fm0, m->mx, m->mv, m->fm,
                                                             gate);
    GC_LAUNCH_CHECK("ms restore");
    return GC_OK;
}

// k substeps, from the current state or (from_saved) from the saved start
// state; with a gate (the surrogate's device-decided fallback) every kernel
// runs only when *gate != 0 and the exact trajectory's residual is not formed
int plain_substeps(gc_engine* e, MembraneSub* m, bool record, long row0, const int* gate, bool from_saved)
{
    hipStream_t s = e->s;
    const int nb = blocks_for(m->nm, BS);
    const float dts = e->hreps[0].P.dt / (float)m->k, inv_m = e->hreps[0].P.inv_m;
    const bool keep = !gate && m->k == KSUB;   // positions after substeps 6 and 7 for the residual
    const bool cell = use_cell_kernel(e) && !record && e->step != -1;
    int st;
    if (cell) {
        // all substeps of a cell in one workgroup, which also forms d2x
        CellArgs ca = cell_args(e, e->step, 1);
        ca.gate = gate;
        SubArgs sa{m->k, dts, inv_m,
                   from_saved ? m->mx0 : m->mx, from_saved ? m->mv0 : m->mv, from_saved ? m->fm0 : m->fm,
                   m->mx, m->mv, m->fe, m->fm, m->mcoor,
                   keep ? m->mx6 : nullptr, keep ? m->mx7 : nullptr, keep ? m->d2x : nullptr};
        k_cell_substeps<<<e->ncell, GC_CELL_BS, 0, s>>>(m->Pm, ca, ms_defargs(e, m), sa);
        GC_LAUNCH_CHECK("membrane substeps (per cell)");
    } else {
        if (from_saved && (st = restore_start(e, m, gate))) return st;
        for (int q = 1; q <= m->k; ++q) {
            k_ms_kick_drift<<<nb, BS, 0, s>>>(m->nm, dts, inv_m, m->mx, m->mv, m->fe, m->fm, m->mcoor, gate);
            if ((st = ms_force(e, m, gate))) return st;
            k_ms_kick<<<nb, BS, 0, s>>>(m->nm, dts, inv_m, m->mv, m->fe, m->fm, gate);
            if (record)
                k_ms_target<<<nb, BS, 0, s>>>(m->nm, q, dts, m->mx, m->mx0, m->mv0, m->frame, m->lref,
                                               m->rec_y + (size_t)NO * row0);
            GC_LAUNCH_CHECK("membrane substep");
            if (keep && (q == KSUB - 2 || q == KSUB - 1))
                k_ms_copy_d<<<blocks_for(3 * m->nm, BS), BS, 0, s>>>(3 * m->nm, m->mx,
                                                                     q == KSUB - 2 ? m->mx6 : m->mx7);
        }
        if (keep) k_ms_d2x<<<nb, BS, 0, s>>>(m->nm, m->mx, m->mx7, m->mx6, m->d2x);
    }
    if (keep) {  // the residual of the exact trajectory (scale for the threshold)
        k_ms_finish<<<nb, BS, 0, s>>>(m->nm, 0, dts, inv_m, m->mv, m->fe, m->fm, m->d2x, m->lref, m->part);
        k_ms_res_final<<<1, RES_BS, 0, s>>>(nb, m->nm, m->part, m->res + 1, 0.0f, nullptr);
        GC_LAUNCH_CHECK("membrane plain residual");
    }
    return GC_OK;
}

}  // namespace

namespace gc {

int ms_reset(gc_engine* e)
{
    MembraneSub* m = e->ms;
    const int nb = blocks_for(m->nm, BS);
    k_ms_init<<<nb, BS, 0, e->s>>>(m->nm, (int)e->consts.nCellStart, e->N, e->d_curpos, e->d_pos64[e->cur],
                                   e->d_nt[e->cur], e->d_vout, e->d_bbox, per3(e), m->mx, m->mv, m->mcoor);
    GC_LAUNCH_CHECK("membrane substep init");
    return ms_force(e, m);
}

bool ms_graph_ok(const gc_engine* e)
{
    const MembraneSub* m = e->ms;
    if (m->mode == 1) return m->rec_x == nullptr;
    return use_cell_kernel(e);
}

int ms_step(gc_engine* e)
{
    MembraneSub* m = e->ms;
    hipStream_t s = e->s;
    const int nm = m->nm, nc = (int)e->consts.nCellStart;
    const int nb = blocks_for(nm, BS);
    const bool record = m->mode == 1 && m->rec_x && m->k == KSUB && m->rec_count + nm <= m->rec_cap;
    // the start state is kept for the surrogate (its prediction and
    // fallback) and for recorded targets
    const bool save = m->mode == 2 || record;
    k_ms_fext<<<nb, BS, 0, s>>>(nm, nc, e->d_curpos, e->d_force, m->fe, m->mx, m->mv, m->fm,
                                save ? m->mx0 : nullptr, m->mv0, m->fm0);
    GC_LAUNCH_CHECK("membrane fluid force");
    int st;
    if (m->mode == 1) {
        if (record) {
            k_ms_features<<<nb, BS, 0, s>>>(feat_args(e, m, false));
            GC_LAUNCH_CHECK("membrane features");
            if ((st = hip_check(hipMemcpyAsync(m->rec_x + (size_t)NF * m->rec_count, m->feat,
                                               sizeof(float) * NF * (size_t)nm, hipMemcpyDeviceToDevice, s),
                                "record features")))
                return st;
        }
        if ((st = plain_substeps(e, m, record, m->rec_count, nullptr, false))) return st;
        if (record) m->rec_count += nm;
    } else {
        const float dts = e->hreps[0].P.dt / (float)KSUB, inv_m = e->hreps[0].P.inv_m;
        k_ms_features<<<nb, BS, 0, s>>>(feat_args(e, m, true));
        GC_LAUNCH_CHECK("membrane features");
        if ((st = mlp_forward(m->mlp, m->feat, m->pred, nm, s))) return st;
        k_ms_apply<<<nb, BS, 0, s>>>(nm, dts, m->pred, m->have_norm ? m->norm : nullptr, m->frame, m->lref,
                                     m->mx0, m->mv0, m->mx, m->mv, m->mcoor, m->d2x);
        GC_LAUNCH_CHECK("membrane surrogate apply");
        if ((st = ms_force(e, m, nullptr, false))) return st;
        const bool dev = use_cell_kernel(e);
        // with the per-cell kernel the fallback is decided on the device: the
        // residual kernel sets the flag that gates the plain substeps from the
        // saved start (no host read; the step can be replayed from a graph)
        k_ms_finish<<<nb, BS, 0, s>>>(nm, 1, dts, inv_m, m->mv, m->fe, m->fm, m->d2x, m->lref, m->part);
        k_ms_res_final<<<1, RES_BS, 0, s>>>(nb, nm, m->part, m->res, dev ? m->thr : 0.0f, dev ? m->flag : nullptr);
        GC_LAUNCH_CHECK("membrane residual");
        if (dev) {
            if ((st = plain_substeps(e, m, false, 0, m->flag, true))) return st;
        } else {
            float r = 0.f;
            if ((st = hip_check(hipMemcpyAsync(&r, m->res, sizeof(float), hipMemcpyDeviceToHost, s), "residual D2H")) ||
                (st = hip_check(hipStreamSynchronize(s), "residual sync")))
                return st;
            if (!(r <= m->thr)) {   // also catches NaN
                m->fallbacks += 1;
                if ((st = plain_substeps(e, m, false, 0, nullptr, true))) return st;
            }
        }
    }
    k_ms_end<<<nb, BS, 0, s>>>(nm, nc, e->N, e->d_curpos, m->mx, m->mv, m->fm, e->d_bbox, per3(e),
                               e->d_pos64[e->cur], e->d_nt[e->cur], e->d_coor, e->d_vout, e->d_defforce, m->flag,
                               e->skin_mode() ? e->d_woff : nullptr, e->d_rbpos,
                               e->skin_mode() ? e->d_bmax + (size_t)DSTAT_N * ((e->N + 63) / 64) : nullptr);
    GC_LAUNCH_CHECK("membrane write-back");
    return GC_OK;
}

}  // namespace gc

namespace {

// 1-ring of every membrane vertex, counter-clockwise about the outward
// normal of the (consistently oriented) triangles, starting at the smallest
// neighbour index
int build_rings(int nm, const std::vector<tint3>& tri, std::vector<int>& ring)
{
    std::vector<std::map<int, int>> next(nm);
    for (const tint3& t : tri) {
        int c[3] = {t.x - 1, t.y - 1, t.z - 1};
        for (int k = 0; k < 3; ++k) {
            int v = c[k], u = c[(k + 1) % 3], w = c[(k + 2) % 3];
            if (v < 0 || v >= nm) return set_error(GC_ERR_TOPOLOGY, "membrane ring: vertex id out of range");
            next[v][u] = w;
        }
    }
    ring.assign((size_t)nm * 6, -1);
    for (int v = 0; v < nm; ++v) {
        const auto& nx = next[v];
        const int deg = (int)nx.size();
        if (deg < 3 || deg > 6)
            return set_error(GC_ERR_TOPOLOGY, "membrane surrogate needs vertices of 3 to 6 neighbours");
        int u = nx.begin()->first;
        for (int k = 0; k < deg; ++k) {
            ring[6 * (size_t)v + k] = u;
            auto it = nx.find(u);
            if (it == nx.end()) return set_error(GC_ERR_TOPOLOGY, "membrane ring is not closed");
            u = it->second;
        }
        if (u != nx.begin()->first) return set_error(GC_ERR_TOPOLOGY, "membrane ring is not closed");
    }
    return GC_OK;
}

}  // namespace

extern "C" int gc_EngineSetMembraneMode(gc_engine* e, int mode, int k, gc_mlp* mlp, const float* norm,
                                        float threshold)
{
    if (!e || mode < 0 || mode > 2) return set_error(GC_ERR_INVALID, "SetMembraneMode: bad arguments");
    if (mode == 0) {
        if (!e->ms) return GC_OK;
        MembraneSub* m = e->ms;
        k_ms_addforce<<<blocks_for(m->nm, BS), BS, 0, e->s>>>(m->nm, (int)e->consts.nCellStart, e->d_curpos,
                                                               m->fm, e->d_force);
        int st = hip_check(hipGetLastError(), "SetMembraneMode");
        int st2 = hip_check(hipStreamSynchronize(e->s), "SetMembraneMode sync");
        ms_destroy(m);
        e->ms = nullptr;
        return st ? st : st2;
    }
    if (e->nrep != 1) return set_error(GC_ERR_INVALID, "SetMembraneMode: batched engine");
    if (e->ncell == 0) return set_error(GC_ERR_INVALID, "SetMembraneMode: no membrane");
    if (k < 1 || (mode == 2 && k != KSUB))
        return set_error(GC_ERR_INVALID, "SetMembraneMode: k must be >= 1 (8 for the surrogate)");
    if (mode == 2 && (!mlp || !(threshold > 0)))
        return set_error(GC_ERR_INVALID, "SetMembraneMode: the surrogate needs a network and a threshold > 0");
    // every argument is checked before the engine is touched: a rejected call
    // leaves the engine as it was
    if (norm) {
        for (int c = 0; c < NF; ++c)
            if (!(norm[NF + c] > 0)) return set_error(GC_ERR_INVALID, "SetMembraneMode: feature std must be > 0");
        for (int c = 0; c < NO; ++c)
            if (!(norm[2 * NF + NO + c] > 0)) return set_error(GC_ERR_INVALID, "SetMembraneMode: target std must be > 0");
    }
    int st = hip_check(hipStreamSynchronize(e->s), "SetMembraneMode sync");
    if (st) return st;
    if (e->gexec) { (void)hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
    MembraneSub* m = e->ms;
    const bool fresh = m == nullptr;
    if (fresh) {
        // built aside and installed only once it is complete
        m = new MembraneSub();
        const size_t nm = (size_t)e->nm;
        m->nm = e->nm;
        if ((st = m->alloc(&m->mx, 3 * nm)) || (st = m->alloc(&m->mx0, 3 * nm)) ||
            (st = m->alloc(&m->mx6, 3 * nm)) || (st = m->alloc(&m->mx7, 3 * nm)) || (st = m->alloc(&m->mv, nm)) ||
            (st = m->alloc(&m->mv0, nm)) || (st = m->alloc(&m->fm, nm)) || (st = m->alloc(&m->fm0, nm)) ||
            (st = m->alloc(&m->fe, nm)) || (st = m->alloc(&m->mcoor, nm)) || (st = m->alloc(&m->d2x, nm)) ||
            (st = m->alloc(&m->mid, nm)) || (st = m->alloc(&m->mnt, nm)) || (st = m->alloc(&m->ring, 6 * nm)) ||
            (st = m->alloc(&m->lref, nm)) || (st = m->alloc(&m->frame, 9 * nm)) ||
            (st = m->alloc(&m->feat, (size_t)NF * nm)) || (st = m->alloc(&m->pred, (size_t)NO * nm)) ||
            (st = m->alloc(&m->norm, 2 * (NF + NO))) || (st = m->alloc(&m->part, (size_t)blocks_for(m->nm, BS))) ||
            (st = m->alloc(&m->res, 2)) || (st = m->alloc(&m->flag, 4))) {
            ms_destroy(m);
            return st;
        }
        // host-side topology: rings from the triangles, rest edge length per vertex
        std::vector<tint3> tri(e->ntri);
        std::vector<int6> edge(e->nedge);
        std::vector<float> l0(e->nedge);
        if ((st = hip_check(hipMemcpy(tri.data(), e->d_tri, sizeof(tint3) * tri.size(), hipMemcpyDeviceToHost), "ms topo")) ||
            (st = hip_check(hipMemcpy(edge.data(), e->d_edge, sizeof(int6) * edge.size(), hipMemcpyDeviceToHost), "ms topo")) ||
            (st = hip_check(hipMemcpy(l0.data(), e->d_l0, sizeof(float) * l0.size(), hipMemcpyDeviceToHost), "ms topo"))) {
            ms_destroy(m);
            return st;
        }
        std::vector<int> ring;
        if ((st = build_rings(m->nm, tri, ring))) { ms_destroy(m); return st; }
        std::vector<double> lsum(nm, 0.0);
        std::vector<int> lcnt(nm, 0);
        for (size_t q = 0; q < edge.size(); ++q)
            for (int end : {edge[q].x - 1, edge[q].y - 1}) {
                lsum[end] += l0[q];
                lcnt[end] += 1;
            }
        std::vector<float> lref(nm);
        std::vector<int> mid(nm);
        std::vector<tint3> zero(nm, tint3{0, 0, 0});
        for (size_t v = 0; v < nm; ++v) {
            lref[v] = lcnt[v] ? (float)(lsum[v] / lcnt[v]) : 1.0f;
            mid[v] = (int)v;
        }
        hipStream_t s = e->s;
        if ((st = upload(m->ring, ring.data(), ring.size(), s)) || (st = upload(m->lref, lref.data(), nm, s)) ||
            (st = upload(m->mid, mid.data(), nm, s)) || (st = upload(m->mnt, zero.data(), nm, s)) ||
            (st = hip_check(hipMemsetAsync(m->flag, 0, 4 * sizeof(int), s), "ms flags"))) {
            ms_destroy(m);
            return st;
        }
        // the membrane as a system of its own: no fluid ids, no periodic axes
        m->Pm = e->hreps[0].P;
        m->Pm.nb2 = m->Pm.nf = m->Pm.nc = 0;
        m->Pm.N = m->nm;
        m->Pm.per[0] = m->Pm.per[1] = m->Pm.per[2] = 0;
    }
    if (norm && (st = upload(m->norm, norm, 2 * (NF + NO), e->s))) {
        if (fresh) ms_destroy(m);
        return st;
    }
    m->mode = mode;
    m->k = k;
    m->mlp = mlp;
    m->thr = threshold;
    m->have_norm = norm != nullptr;
    if (fresh) {
        e->ms = m;
        if ((st = ms_reset(e))) {
            (void)hipStreamSynchronize(e->s);
            ms_destroy(m);
            e->ms = nullptr;
            return st;
        }
    }
    return hip_check(hipStreamSynchronize(e->s), "SetMembraneMode sync");
}

extern "C" int gc_EngineRecordSamples(gc_engine* e, float* x, float* y, int capacity)
{
    if (!e || !e->ms) return set_error(GC_ERR_INVALID, "RecordSamples: engine not in mode 1");
    if (x && (!y || capacity <= 0)) return set_error(GC_ERR_INVALID, "RecordSamples: bad buffers");
    MembraneSub* m = e->ms;
    m->rec_x = x;
    m->rec_y = x ? y : nullptr;
    m->rec_cap = x ? capacity : 0;
    m->rec_count = 0;
    return GC_OK;
}

extern "C" int gc_EngineMembraneStats(gc_engine* e, gc_membrane_stats* out)
{
    if (!e || !out) return set_error(GC_ERR_INVALID, "MembraneStats: null");
    memset(out, 0, sizeof(*out));
    if (!e->ms) return GC_OK;
    out->samples = e->ms->rec_count;
    float r[2] = {0.f, 0.f};
    int f[3] = {0, 0, 0};
    int st = hip_check(hipStreamSynchronize(e->s), "MembraneStats sync");
    if (!st) st = hip_check(hipMemcpy(r, e->ms->res, sizeof(r), hipMemcpyDeviceToHost), "MembraneStats");
    if (!st) st = hip_check(hipMemcpy(f, e->ms->flag, sizeof(f), hipMemcpyDeviceToHost), "MembraneStats");
    out->steps = f[2];
    out->fallbacks = e->ms->fallbacks + f[1];
    out->last_residual = r[0];
    out->plain_residual = r[1];
    return st;
}
