// gc_engine.hip -- the engine: library-owned state and the step of
// Particle::UpdatePar (src/Particle.cu:1310-1329) as a fixed launch sequence
// on one stream, capturable into a HIP graph.
//
// Differences from the one-call-per-entry-point sequence, all internal:
//   * positions are accumulated in FP64 (pos64); the FP32 coordinates the
//     neighbour kernels read are rounded from it once per step;
//   * integrate + periodic wrap + cube key are one kernel;
//   * the cell list is a counting sort keyed by (cube, id) with the cube
//     offsets of the scan used directly as BeginEndCube (no host sync);
//   * the sorted gather and the velocity predictor are one kernel;
//   * the noise counter lives in device memory, so graph replays advance it.
#include "gc_kernels.h"

#include <math.h>
#include <string.h>
#include <vector>

using namespace gc;

namespace {

constexpr int BS = GC_BS;

__global__ __launch_bounds__(BS) void k_e_integrate(int n, GcParams P, int move,
    double* __restrict__ pos64, const tfloat4* __restrict__ velrhop,
    const tfloat3* __restrict__ force, const int* __restrict__ sortpart, tint3* __restrict__ nthro,
    const float* __restrict__ bbox, const float* __restrict__ invw, const int* __restrict__ ncube,
    int* __restrict__ key)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int id = sortpart[p];
    double x = pos64[3 * p], y = pos64[3 * p + 1], z = pos64[3 * p + 2];
    if (move && id >= P.nf) {
        double dt = P.dt, a = 0.5 * (double)P.dt * (double)P.dt * (double)P.inv_m;
        tfloat4 v = velrhop[p];
        tfloat3 f = force[p];
        x += dt * v.x + a * f.x;
        y += dt * v.y + a * f.y;
        z += dt * v.z + a * f.z;
        double u[3] = {x, y, z};
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
        pos64[3 * p] = x; pos64[3 * p + 1] = y; pos64[3 * p + 2] = z;
    }
    int nx = ncube[0], ny = ncube[1], nz = ncube[2];
    int cx = cube_coord((float)x, bbox[0], P.off[0], invw[0], nx);
    int cy = cube_coord((float)y, bbox[2], P.off[1], invw[1], ny);
    int cz = cube_coord((float)z, bbox[4], P.off[2], invw[2], nz);
    key[p] = cx + nx * (cy + ny * cz);
}

// New slot s takes the particle whose old slot is curpos[id]; applies the
// Groot-Warren predictor (PredictFluidVel) on the way.
__global__ __launch_bounds__(BS) void k_e_gather(int n, GcParams P, int predict,
    const int* __restrict__ sortpart_new, int* __restrict__ curpos, const int* __restrict__ key_new,
    const int* __restrict__ ncube, const double* __restrict__ pos_old, double* __restrict__ pos_new,
    const tfloat4* __restrict__ vr_old, tfloat4* __restrict__ vr_new,
    const tint3* __restrict__ nt_old, tint3* __restrict__ nt_new, const tfloat3* __restrict__ force,
    tfloat3* __restrict__ velhalf, tfloat3* __restrict__ coor, tint3* __restrict__ cc)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int id = sortpart_new[s];
    int o = curpos[id];
    double x = pos_old[3 * o], y = pos_old[3 * o + 1], z = pos_old[3 * o + 2];
    pos_new[3 * s] = x; pos_new[3 * s + 1] = y; pos_new[3 * s + 2] = z;
    coor[s] = tfloat3{(float)x, (float)y, (float)z};
    int k = key_new[s], nx = ncube[0], ny = ncube[1];
    cc[s] = tint3{k % nx, (k / nx) % ny, k / (nx * ny)};
    nt_new[s] = nt_old[o];
    tfloat4 v = vr_old[o];
    tfloat3 f = force[o];
    if (predict && id >= P.nf) {
        float h = 0.5f * P.dt * P.inv_m, l = P.lam * P.dt * P.inv_m;
        velhalf[s] = tfloat3{v.x + h * f.x, v.y + h * f.y, v.z + h * f.z};
        v.x += l * f.x; v.y += l * f.y; v.z += l * f.z;
    } else {
        velhalf[s] = tfloat3{v.x, v.y, v.z};
    }
    vr_new[s] = v;
    curpos[id] = s;
}

__global__ void k_e_begin_end(int K, const int* __restrict__ offs, tint2* __restrict__ be)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < K) be[k] = tint2{offs[k], offs[k + 1]};
}

__global__ void k_e_bump(unsigned long long* ctr) { *ctr += 1ull; }

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

    // grid
    float *d_bbox, *d_outer, *d_invw; int* d_ncube;
    // slot state (double-buffered where the gather reads and writes)
    double* d_pos64[2]; tfloat4* d_vr[2]; tint3* d_nt[2]; int* d_sp[2];
    tfloat3 *d_coor, *d_velhalf, *d_force; float* d_press; tint3* d_cc;
    int *d_key, *d_cubepart, *d_curpos;
    tint2* d_be;
    int *d_cnt, *d_offs, *d_rank, *d_tk, *d_tv, *d_scanws;
    unsigned long long* d_ctr;
    // membrane
    tint3* d_tri; int6* d_edge; int* d_celltab; tint2* d_betri;
    float *d_mu0, *d_kc, *d_kd, *d_ka, *d_kv;
    float *d_l0, *d_lmax, *d_kb, *d_mexp, *d_p, *d_kp, *d_th0;
    tfloat3 *d_trinor, *d_tricen; float *d_triarea, *d_reftriarea;
    float *d_refarea_para, *d_refvol_para, *d_area_para, *d_vol_para, *d_area, *d_vol, *d_pararea;
    tfloat3 *d_defforce, *d_aggforce;
    int *d_es, *d_ee, *d_ts, *d_te;
    // graph
    hipGraphExec_t gexec = nullptr;
    int g_driveon = -1;

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

// Cell list: keys from the (optionally advanced) FP64 positions, counting
// sort by (cube, id), cube offsets -> BeginEndCube, sorted gather.
int launch_build(gc_engine* e, const GcParams& P, int move, int predict)
{
    const int N = e->N, c = e->cur, nx = 1 - c;
    hipStream_t s = e->s;
    k_e_integrate<<<blocks_for(N, BS), BS, 0, s>>>(N, P, move, e->d_pos64[c], e->d_vr[c],
        e->d_force, e->d_sp[c], e->d_nt[c], e->d_bbox, e->d_invw, e->d_ncube, e->d_key);
    GC_LAUNCH_CHECK("engine integrate");
    int st = counting_sort(N, e->K, e->d_key, e->d_sp[c], e->d_cubepart, e->d_sp[nx], e->d_cnt,
                           e->d_offs, e->d_rank, e->d_tk, e->d_tv, s, e->d_scanws);
    if (st) return st;
    k_e_begin_end<<<blocks_for(e->K, 256), 256, 0, s>>>(e->K, e->d_offs, e->d_be);
    GC_LAUNCH_CHECK("engine begin/end");
    k_e_gather<<<blocks_for(N, BS), BS, 0, s>>>(N, P, predict, e->d_sp[nx], e->d_curpos,
        e->d_cubepart, e->d_ncube, e->d_pos64[c], e->d_pos64[nx], e->d_vr[c], e->d_vr[nx],
        e->d_nt[c], e->d_nt[nx], e->d_force, e->d_velhalf, e->d_coor, e->d_cc);
    GC_LAUNCH_CHECK("engine gather");
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
    if (!pr) return;
    (void)hipEventRecord(pr->ev[pr->n], s);
    pr->name[pr->n] = name;
    pr->n++;
}

// One Particle::UpdatePar.
int launch_step(gc_engine* e, int driveon, Prof* pr)
{
    GcParams P = engine_params(e);
    const int N = e->N;
    hipStream_t s = e->s;
    mark(pr, s, "start");
    int st = launch_build(e, P, 1, 1);
    if (st) return st;
    mark(pr, s, "neighbour build+predict");
    const int c = e->cur;
    k_density<<<blocks_for(N, BS), BS, 0, s>>>(N, P, e->d_coor, e->d_vr[c], e->d_cc, e->d_ncube,
                                               e->d_outer, e->d_sp[c], e->d_be);
    GC_LAUNCH_CHECK("engine density");
    mark(pr, s, "density");
    if (P.nf > P.nb2) {
        k_bound2_vel<<<blocks_for(N, BS), BS, 0, s>>>(N, P, e->d_coor, e->d_vr[c], e->d_cc,
                                                      e->d_ncube, e->d_outer, e->d_sp[c], e->d_be);
        GC_LAUNCH_CHECK("engine bound2 vel");
    }
    mark(pr, s, "bound2 velocity");
    k_eos<<<blocks_for(N, BS), BS, 0, s>>>(N, P, e->d_sp[c], e->d_vr[c], e->d_press);
    GC_LAUNCH_CHECK("engine eos");
    mark(pr, s, "eos");
    k_sdpd_force<<<blocks_for(N, BS), BS, 0, s>>>(N, P, e->d_coor, e->d_vr[c], e->d_cc, e->d_ncube,
                                                  e->d_sp[c], e->d_outer, e->d_be, e->d_force,
                                                  e->d_press, e->d_ctr);
    GC_LAUNCH_CHECK("engine sdpd force");
    k_e_bump<<<1, 1, 0, s>>>(e->d_ctr);
    mark(pr, s, "sdpd force");
    if (P.nb2 > 0) {
        k_rep_force<<<blocks_for(N, BS), BS, 0, s>>>(N, P, e->d_coor, e->d_cc, e->d_ncube,
                                                     e->d_sp[c], e->d_outer, e->d_be, e->d_force);
        GC_LAUNCH_CHECK("engine rep force");
    }
    if (driveon) {
        k_ext_force<<<blocks_for(N, BS), BS, 0, s>>>(N, P, e->d_sp[c], e->d_coor, e->d_force);
        GC_LAUNCH_CHECK("engine ext force");
    }
    mark(pr, s, "wall+body force");
    if (e->ncell > 0) {
        k_geo_cell<<<e->ncell, BS, 0, s>>>(P, e->step, e->d_bbox, e->d_curpos, e->d_tri, e->d_betri,
            e->d_nt[c], e->d_coor, e->d_trinor, e->d_triarea, e->d_reftriarea, e->d_tricen,
            e->d_refarea_para, e->d_refvol_para, e->d_area_para, e->d_vol_para, e->d_area, e->d_vol);
        GC_LAUNCH_CHECK("engine geo");
        mark(pr, s, "membrane geometry");
        DefArgs a;
        a.edge = e->d_edge; a.tri = e->d_tri; a.coor = e->d_coor; a.nthro = e->d_nt[c];
        a.curpos = e->d_curpos; a.lmax = e->d_lmax; a.mexp = e->d_mexp; a.p = e->d_p;
        a.kp = e->d_kp; a.th0 = e->d_th0; a.kb = e->d_kb; a.reftriarea = e->d_reftriarea;
        a.kd = e->d_kd; a.ka = e->d_ka; a.Ac = e->d_area_para; a.A0c = e->d_refarea_para;
        a.kv = e->d_kv; a.Vc = e->d_vol_para; a.V0c = e->d_refvol_para;
        a.es = e->d_es; a.ee = e->d_ee; a.ts = e->d_ts; a.te = e->d_te;
        a.defforce = e->d_defforce; a.force = e->d_force;
        k_def_force<<<blocks_for(e->nm, BS), BS, 0, s>>>(P, e->nm, e->d_bbox, a);
        GC_LAUNCH_CHECK("engine def force");
        mark(pr, s, "membrane force");
        if (P.De != 0.0f) {
            k_agg_force<<<blocks_for(N, BS), BS, 0, s>>>(N, P, e->d_coor, e->d_aggforce, e->d_cc,
                e->d_ncube, e->d_sp[c], e->d_celltab, e->d_outer, e->d_be, e->d_force);
            GC_LAUNCH_CHECK("engine agg force");
        }
        mark(pr, s, "morse");
    }
    k_correct<<<blocks_for(N, BS), BS, 0, s>>>(N, P, e->d_sp[c], e->d_force, e->d_velhalf, e->d_vr[c]);
    GC_LAUNCH_CHECK("engine correct");
    mark(pr, s, "corrector");
    if (P.nf > P.nb2) {
        k_bound2_vel<<<blocks_for(N, BS), BS, 0, s>>>(N, P, e->d_coor, e->d_vr[c], e->d_cc,
                                                      e->d_ncube, e->d_outer, e->d_sp[c], e->d_be);
        GC_LAUNCH_CHECK("engine bound2 vel 2");
    }
    mark(pr, s, "bound2 velocity 2");
    e->step += 1;
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
    // validate the parameter block the same way the entry points do
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
    if ((st = e->alloc(&e->d_bbox, 6)) || (st = e->alloc(&e->d_outer, 3)) ||
        (st = e->alloc(&e->d_invw, 3)) || (st = e->alloc(&e->d_ncube, 3)))
        return fail(st);
    for (int b = 0; b < 2; ++b)
        if ((st = e->alloc(&e->d_pos64[b], 3 * N)) || (st = e->alloc(&e->d_vr[b], N)) ||
            (st = e->alloc(&e->d_nt[b], N)) || (st = e->alloc(&e->d_sp[b], N)))
            return fail(st);
    if ((st = e->alloc(&e->d_coor, N)) || (st = e->alloc(&e->d_velhalf, N)) ||
        (st = e->alloc(&e->d_force, N)) || (st = e->alloc(&e->d_press, N)) ||
        (st = e->alloc(&e->d_cc, N)) || (st = e->alloc(&e->d_key, N)) ||
        (st = e->alloc(&e->d_cubepart, N)) || (st = e->alloc(&e->d_curpos, N)) ||
        (st = e->alloc(&e->d_be, K)) || (st = e->alloc(&e->d_cnt, K + 1)) ||
        (st = e->alloc(&e->d_offs, K + 1)) || (st = e->alloc(&e->d_rank, N)) ||
        (st = e->alloc(&e->d_tk, N)) || (st = e->alloc(&e->d_tv, N)) ||
        (st = e->alloc(&e->d_scanws, scan_ws_ints(e->K))) || (st = e->alloc(&e->d_ctr, 1)))
        return fail(st);

    hipStream_t s = e->s;
    // initial state in id order (Particle::SetParticle, Particle.cu:485-546)
    {
        std::vector<double> pos(3 * N);
        std::vector<tfloat4> vr(N);
        std::vector<tfloat3> vh(N), f(N, tfloat3{0.f, 0.f, 0.f});
        std::vector<tint3> nt(N, tint3{0, 0, 0});
        std::vector<int> idn(N);
        for (size_t i = 0; i < N; ++i) {
            pos[3 * i] = d->coor[i].x; pos[3 * i + 1] = d->coor[i].y; pos[3 * i + 2] = d->coor[i].z;
            tfloat3 v = d->vel ? d->vel[i] : tfloat3{0.f, 0.f, 0.f};
            vr[i] = tfloat4{v.x, v.y, v.z, e->consts.DenInit};
            vh[i] = v;
            idn[i] = (int)i;
        }
        unsigned long long ctr = e->ext.noise_counter;
        if ((st = upload(e->d_bbox, e->bbox, 6, s)) || (st = upload(e->d_outer, e->outer, 3, s)) ||
            (st = upload(e->d_invw, e->invw, 3, s)) || (st = upload(e->d_ncube, e->ncube, 3, s)) ||
            (st = upload(e->d_pos64[0], pos.data(), 3 * N, s)) || (st = upload(e->d_vr[0], vr.data(), N, s)) ||
            (st = upload(e->d_velhalf, vh.data(), N, s)) || (st = upload(e->d_force, f.data(), N, s)) ||
            (st = upload(e->d_nt[0], nt.data(), N, s)) || (st = upload(e->d_sp[0], idn.data(), N, s)) ||
            (st = upload(e->d_curpos, idn.data(), N, s)) || (st = upload(e->d_ctr, &ctr, 1, s)))
            return fail(st);
        if ((st = hip_check(hipStreamSynchronize(s), "engine upload sync"))) return fail(st);
    }
    GcParams P = engine_params(e);
    // first neighbour build (Particle::ConCubeList at init, Particle.cu:570)
    if ((st = launch_build(e, P, 0, 0))) return fail(st);
    k_eos<<<blocks_for(e->N, BS), BS, 0, s>>>(e->N, P, e->d_sp[e->cur], e->d_vr[e->cur], e->d_press);

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
            (st = e->alloc(&e->d_defforce, M)) || (st = e->alloc(&e->d_aggforce, M)))
            return fail(st);
        std::vector<int> es, ee, ts, te;
        if ((st = build_incidence(e->nm, e->nedge, d->edge, e->ntri, d->tri, es, ee, ts, te)))
            return fail(st);
        for (int c = 0; c < e->ncell; ++c)
            if (d->beginendtri[c].x < 0 || d->beginendtri[c].y >= e->ntri ||
                d->beginendtri[c].x > d->beginendtri[c].y)
                return fail(set_error(GC_ERR_TOPOLOGY, "EngineCreate: bad BeginEndTri"));
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
    // A step alternates the double buffers, so one graph covers two steps.
    if (!e->gexec || e->g_driveon != driveon || (e->cur != 0)) {
        if (e->cur != 0) {
            int st = launch_step(e, driveon, nullptr);
            if (st) return st;
            if (--nsteps == 0) return GC_OK;
        }
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
    (void)hipStreamSynchronize(e->s);
    (void)hipMemcpy(&c, e->d_ctr, sizeof(c), hipMemcpyDeviceToHost);
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
    hipStream_t s = e->s;
    const int N = e->N, c = e->cur;
    int st = hip_check(hipStreamSynchronize(s), "GetState sync");
    if (st) return st;
    std::vector<int> sp(N);
    std::vector<double> pos(3 * (size_t)N);
    std::vector<tfloat4> vr(N);
    std::vector<tfloat3> vh(N), f(N), co(N);
    std::vector<float> pr(N);
    std::vector<tint3> nt(N);
    auto dl = [&](void* h, const void* dv, size_t b) {
        return hip_check(hipMemcpy(h, dv, b, hipMemcpyDeviceToHost), "GetState D2H");
    };
    if ((st = dl(sp.data(), e->d_sp[c], sizeof(int) * N)) ||
        (st = dl(pos.data(), e->d_pos64[c], sizeof(double) * 3 * N)) ||
        (st = dl(vr.data(), e->d_vr[c], sizeof(tfloat4) * N)) ||
        (st = dl(vh.data(), e->d_velhalf, sizeof(tfloat3) * N)) ||
        (st = dl(f.data(), e->d_force, sizeof(tfloat3) * N)) ||
        (st = dl(co.data(), e->d_coor, sizeof(tfloat3) * N)) ||
        (st = dl(pr.data(), e->d_press, sizeof(float) * N)) ||
        (st = dl(nt.data(), e->d_nt[c], sizeof(tint3) * N)))
        return st;
    for (int slot = 0; slot < N; ++slot) {
        int id = sp[slot];
        if (o->pos) for (int d = 0; d < 3; ++d) o->pos[3 * (size_t)id + d] = pos[3 * (size_t)slot + d];
        if (o->coor) o->coor[id] = co[slot];
        if (o->velrhop) o->velrhop[id] = vr[slot];
        if (o->velhalf) o->velhalf[id] = vh[slot];
        if (o->force) o->force[id] = f[slot];
        if (o->press) o->press[id] = pr[slot];
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
    // current state in id order, then overwrite with the given fields
    std::vector<double> pos(3 * (size_t)N);
    std::vector<tfloat4> vr(N);
    std::vector<tfloat3> vh(N), f(N);
    std::vector<float> pr(N);
    std::vector<tint3> nt(N);
    gc_host_state cur;
    memset(&cur, 0, sizeof(cur));
    cur.pos = pos.data(); cur.velrhop = vr.data(); cur.velhalf = vh.data(); cur.force = f.data();
    cur.press = pr.data(); cur.nthrobound = nt.data();
    if ((st = gc_EngineGetState(e, &cur))) return st;
    if (in->pos) memcpy(pos.data(), in->pos, sizeof(double) * 3 * N);
    if (in->velrhop) memcpy(vr.data(), in->velrhop, sizeof(tfloat4) * N);
    if (in->velhalf) memcpy(vh.data(), in->velhalf, sizeof(tfloat3) * N);
    if (in->force) memcpy(f.data(), in->force, sizeof(tfloat3) * N);
    if (in->press) memcpy(pr.data(), in->press, sizeof(float) * N);
    if (in->nthrobound) memcpy(nt.data(), in->nthrobound, sizeof(tint3) * N);
    std::vector<int> idn(N);
    for (int i = 0; i < N; ++i) idn[i] = i;
    e->cur = 0;
    hipStream_t s = e->s;
    unsigned long long ctr = noise_counter;
    if ((st = upload(e->d_pos64[0], pos.data(), 3 * (size_t)N, s)) || (st = upload(e->d_vr[0], vr.data(), N, s)) ||
        (st = upload(e->d_velhalf, vh.data(), N, s)) || (st = upload(e->d_force, f.data(), N, s)) ||
        (st = upload(e->d_press, pr.data(), N, s)) || (st = upload(e->d_nt[0], nt.data(), N, s)) ||
        (st = upload(e->d_sp[0], idn.data(), N, s)) || (st = upload(e->d_curpos, idn.data(), N, s)) ||
        (st = upload(e->d_ctr, &ctr, 1, s)))
        return st;
    GcParams P = engine_params(e);
    if ((st = launch_build(e, P, 0, 0))) return st;
    // the gather without predictor copies velrhop into velhalf and does not
    // move force and pressure: put those three into the new slot order
    {
        std::vector<int> sp(N);
        if ((st = hip_check(hipStreamSynchronize(s), "SetState sync2"))) return st;
        if ((st = hip_check(hipMemcpy(sp.data(), e->d_sp[e->cur], sizeof(int) * N, hipMemcpyDeviceToHost), "SetState sp")))
            return st;
        std::vector<tfloat3> vhs(N), fs(N);
        std::vector<float> prs(N);
        for (int slot = 0; slot < N; ++slot) {
            vhs[slot] = vh[sp[slot]]; fs[slot] = f[sp[slot]]; prs[slot] = pr[sp[slot]];
        }
        if ((st = upload(e->d_velhalf, vhs.data(), N, s)) || (st = upload(e->d_press, prs.data(), N, s)) ||
            (st = upload(e->d_force, fs.data(), N, s)))
            return st;
    }
    e->step = step;
    return hip_check(hipStreamSynchronize(s), "SetState sync3");
}

extern "C" int gc_EngineGetCells(gc_engine* e, gc_host_cells* o)
{
    if (!e || !o) return set_error(GC_ERR_INVALID, "EngineGetCells: null");
    int st = hip_check(hipStreamSynchronize(e->s), "GetCells sync");
    if (st) return st;
    const int N = e->N, c = e->cur;
    auto dl = [&](void* h, const void* dv, size_t b) {
        return h ? hip_check(hipMemcpy(h, dv, b, hipMemcpyDeviceToHost), "GetCells D2H") : GC_OK;
    };
    if ((st = dl(o->sortpart, e->d_sp[c], sizeof(int) * N)) ||
        (st = dl(o->cubepart, e->d_cubepart, sizeof(int) * N)) ||
        (st = dl(o->cc, e->d_cc, sizeof(tint3) * N)) ||
        (st = dl(o->curpos, e->d_curpos, sizeof(int) * N)) ||
        (st = dl(o->beginendcube, e->d_be, sizeof(tint2) * e->K)))
        return st;
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
