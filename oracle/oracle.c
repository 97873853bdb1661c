/* oracle.c -- FP64 CPU restatement of the SDPD step (checker only; see
 * oracle.h for what it restates and how each function is pinned).
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).           */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979323846

/* ------------------------------------------------------------------ */
/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random
 * numbers: as easy as 1, 2, 3"): 10 rounds of the 4x32 S-P network with
 * multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments
 * 0x9E3779B9 / 0xBB67AE85.                                              */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* uniform in (0,1): 24 high bits, centred, exactly representable in FP32 */
static double u01(uint32_t x) { return ((double)(x >> 8) + 0.5) * (1.0 / 16777216.0); }

/* Counter = (min id, max id, counter lo, counter hi), key = seed halves;
 * Box-Muller on (u0,u1) and (u2,u3), the fourth normal is unused. */
void orc_pair_normals(uint64_t seed, uint64_t counter, uint32_t ida, uint32_t idb,
                      double n[3])
{
    uint32_t lo = ida < idb ? ida : idb, hi = ida < idb ? idb : ida;
    uint32_t ctr[4] = {lo, hi, (uint32_t)counter, (uint32_t)(counter >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    double r0 = sqrt(-2.0 * log(u01(o[0])));
    double r1 = sqrt(-2.0 * log(u01(o[2])));
    double t0 = 2.0 * PI * u01(o[1]);
    double t1 = 2.0 * PI * u01(o[3]);
    n[0] = r0 * cos(t0);
    n[1] = r0 * sin(t0);
    n[2] = r1 * cos(t1);
}

/* Lucy (1977) 3-D kernel with support h. */
double orc_lucy_w(double r, double h)
{
    double q = r / h;
    if (q >= 1.0) return 0.0;
    double om = 1.0 - q;
    return 105.0 / (16.0 * PI * h * h * h) * (1.0 + 3.0 * q) * om * om * om;
}

double orc_lucy_f(double r, double h)
{
    double q = r / h;
    if (q >= 1.0) return 0.0;
    double om = 1.0 - q;
    return 315.0 / (4.0 * PI * h * h * h * h * h) * om * om;
}

/* ------------------------------------------------------------------ */
/* Cell list with FP32 keys (Particle::ConfigDomain, Particle.cu:548-571;
 * cudiv::PreConList / Sort / CalBeginEndCube, ConfigDomain.cuh:11-13).   */
int orc_config_domain(const float bbox[6], float cut, const int periodic[3],
                      int ncube[3], float outer[3], float invwid[3], int* sizecube)
{
    for (int d = 0; d < 3; ++d) {
        float L = bbox[2 * d + 1] - bbox[2 * d];
        if (!periodic[d]) L = L + 2.0f * cut;
        outer[d] = L;
        ncube[d] = (int)floorf(L / cut);
        if (ncube[d] < 1) return -1;
        invwid[d] = (float)ncube[d] / outer[d];
        if (periodic[d] && ncube[d] < 3) return -1;
    }
    *sizecube = ncube[0] * ncube[1] * ncube[2];
    return 0;
}

void orc_preconlist(int n, float cut, const int periodic[3], const float* coor,
                    const float bbox[6], const float invwid[3], const int ncube[3],
                    int* cc, int* cubepart, const int* sortpart, int* oldpos)
{
    for (int p = 0; p < n; ++p) {
        int c[3];
        for (int d = 0; d < 3; ++d) {
            volatile float off = periodic[d] ? 0.0f : cut;
            volatile float t = coor[3 * p + d] - bbox[2 * d];
            volatile float u = t + off;
            volatile float f = u * invwid[d];
            int k = (int)floorf(f);
            if (k < 0) k = 0;
            if (k > ncube[d] - 1) k = ncube[d] - 1;
            c[d] = k;
        }
        cc[3 * p + 0] = c[0]; cc[3 * p + 1] = c[1]; cc[3 * p + 2] = c[2];
        cubepart[p] = c[0] + ncube[0] * (c[1] + ncube[1] * c[2]);
        if (oldpos) oldpos[sortpart[p]] = p;
    }
}

typedef struct { int k, v; } kv_t;
static int kv_cmp(const void* a, const void* b)
{
    const kv_t* x = (const kv_t*)a; const kv_t* y = (const kv_t*)b;
    if (x->k != y->k) return x->k < y->k ? -1 : 1;
    if (x->v != y->v) return x->v < y->v ? -1 : 1;
    return 0;
}

void orc_sort(int n, int* keys, int* values)
{
    kv_t* a = (kv_t*)malloc(sizeof(kv_t) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) { a[i].k = keys[i]; a[i].v = values[i]; }
    qsort(a, (size_t)n, sizeof(kv_t), kv_cmp);
    for (int i = 0; i < n; ++i) { keys[i] = a[i].k; values[i] = a[i].v; }
    free(a);
}

void orc_begin_end(int n, int sizecube, const int* cubepart, int* beginend)
{
    /* [first, last+1); an empty cube k gets (b, b) with b = #keys < k */
    int p = 0;
    for (int k = 0; k < sizecube; ++k) {
        while (p < n && cubepart[p] < k) ++p;
        int b = p;
        while (p < n && cubepart[p] == k) ++p;
        beginend[2 * k] = b;
        beginend[2 * k + 1] = p;
    }
}

/* ------------------------------------------------------------------ */
struct orc_sys {
    /* id ranges */
    int nb2, nf, nc, N, nm;
    /* parameters */
    double h, m, c1, c2, rho0, lam, eta, zeta, fext, dt, De, mcut, r0, beta, kT;
    uint64_t seed;
    double lo[3], hi[3], L[3];
    int per[3];
    int force_mode; double force_split;
    double wall_split, wall_u_lo, wall_u_hi;
    int noise_on;
    double wlc_ratio, pow_index, rep_scale;
    /* membrane topology (0-based) */
    int ncell, ntri, nedge;
    int *tri, *edge, *celltab, *tb, *te;
    double *mu0, *kc, *kd, *ka, *kv;
    /* reference state and mechanics */
    double *A0t, *A0c, *V0c;
    double *l0, *lmax, *kb, *mexp, *p, *kp, *th0;
    int have_ref, have_mech;
    /* current cell totals (from the last orc_geo) */
    double *Ac, *Vc;
    /* neighbour list */
    int *nl_start, *nl_j;
    size_t nl_cap;
};

static int g_threads = 0;
void orc_set_threads(int nthreads) { g_threads = nthreads; }

static double* dup_f(const float* a, int n)
{
    double* o = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    if (a) for (int i = 0; i < n; ++i) o[i] = a[i];
    return o;
}

orc_sys* orc_create(const Structconst* c, const gc_ext_consts* ext,
                    const float bbox[6], int ncell, int ntri, int nedge,
                    const int* tri, const int* edge, const int* celltab,
                    const int* beginendtri, const float* shearmod,
                    const float* bendmod, const float* hlarea,
                    const float* hgarea, const float* hvol)
{
    orc_sys* s = (orc_sys*)calloc(1, sizeof(orc_sys));
    s->nb2 = (int)c->nBound2Start; s->nf = (int)c->nFluidStart;
    s->nc = (int)c->nCellStart; s->N = (int)c->nTotalPar; s->nm = (int)c->nCellPar;
    s->h = c->CutRadius; s->m = c->Mass; s->c1 = c->SoundSpeed1; s->c2 = c->SoundSpeed2;
    s->rho0 = c->DenInit; s->lam = c->VVsigma; s->eta = c->ShearVis; s->zeta = c->BulkVis;
    s->fext = c->ConstExtForce; s->dt = c->TimeStep; s->De = c->surEnergy;
    s->mcut = c->MaxCut; s->r0 = c->zeroLen; s->beta = c->scaFac; s->kT = c->Temp;
    s->seed = (uint64_t)c->iSeed;
    for (int d = 0; d < 3; ++d) {
        s->lo[d] = bbox[2 * d]; s->hi[d] = bbox[2 * d + 1];
        /* the period is the FP32 difference, as the kernels see it */
        s->L[d] = (double)(bbox[2 * d + 1] - bbox[2 * d]);
        s->per[d] = ext->periodic[d];
    }
    s->force_mode = ext->force_mode; s->force_split = ext->force_split;
    s->wall_split = ext->wall_split; s->wall_u_lo = ext->wall_u_lo; s->wall_u_hi = ext->wall_u_hi;
    s->noise_on = ext->noise_on;
    s->wlc_ratio = ext->wlc_ratio; s->pow_index = ext->pow_index;
    s->rep_scale = ext->rep_scale;
    s->ncell = ncell; s->ntri = ntri; s->nedge = nedge;
    if (ncell > 0) {
        s->tri = (int*)malloc(sizeof(int) * 3 * (size_t)ntri);
        s->edge = (int*)malloc(sizeof(int) * 6 * (size_t)nedge);
        s->celltab = (int*)malloc(sizeof(int) * (size_t)s->nm);
        s->tb = (int*)malloc(sizeof(int) * (size_t)ncell);
        s->te = (int*)malloc(sizeof(int) * (size_t)ncell);
        for (int t = 0; t < 3 * ntri; ++t) s->tri[t] = tri[t] - 1;
        for (int e = 0; e < nedge; ++e)
            for (int k = 0; k < 6; ++k) s->edge[6 * e + k] = edge[6 * e + k] - 1;
        memcpy(s->celltab, celltab, sizeof(int) * (size_t)s->nm);
        for (int q = 0; q < ncell; ++q) { s->tb[q] = beginendtri[2 * q]; s->te[q] = beginendtri[2 * q + 1]; }
        s->mu0 = dup_f(shearmod, nedge); s->kc = dup_f(bendmod, nedge);
        s->kd = dup_f(hlarea, ntri); s->ka = dup_f(hgarea, ntri); s->kv = dup_f(hvol, ntri);
        s->A0t = (double*)calloc((size_t)ntri, sizeof(double));
        s->A0c = (double*)calloc((size_t)ncell, sizeof(double));
        s->V0c = (double*)calloc((size_t)ncell, sizeof(double));
        s->Ac = (double*)calloc((size_t)ncell, sizeof(double));
        s->Vc = (double*)calloc((size_t)ncell, sizeof(double));
        s->l0 = (double*)calloc((size_t)nedge, sizeof(double));
        s->lmax = (double*)calloc((size_t)nedge, sizeof(double));
        s->kb = (double*)calloc((size_t)nedge, sizeof(double));
        s->mexp = (double*)calloc((size_t)nedge, sizeof(double));
        s->p = (double*)calloc((size_t)nedge, sizeof(double));
        s->kp = (double*)calloc((size_t)nedge, sizeof(double));
        s->th0 = (double*)calloc((size_t)nedge, sizeof(double));
    }
    s->nl_start = (int*)calloc((size_t)s->N + 1, sizeof(int));
    return s;
}

void orc_destroy(orc_sys* s)
{
    if (!s) return;
    free(s->tri); free(s->edge); free(s->celltab); free(s->tb); free(s->te);
    free(s->mu0); free(s->kc); free(s->kd); free(s->ka); free(s->kv);
    free(s->A0t); free(s->A0c); free(s->V0c); free(s->Ac); free(s->Vc);
    free(s->l0); free(s->lmax); free(s->kb); free(s->mexp); free(s->p);
    free(s->kp); free(s->th0);
    free(s->nl_start); free(s->nl_j);
    free(s);
}

/* minimum-image difference a - b on periodic axes */
static void rel_vec(const orc_sys* s, const double* a, const double* b, double* r)
{
    for (int d = 0; d < 3; ++d) {
        double x = a[d] - b[d];
        if (s->per[d]) {
            if (x > 0.5 * s->L[d]) x -= s->L[d];
            else if (x < -0.5 * s->L[d]) x += s->L[d];
        }
        r[d] = x;
    }
}

/* Neighbour list of every pair closer than h (all particle types),
 * neighbours of i in ascending scan order.  Brute force for small systems,
 * otherwise a bin grid of edge >= h. */
static void build_nlist(orc_sys* s, const double* pos)
{
    const int N = s->N;
    const double h2 = s->h * s->h;
    int nb[3]; double blo[3], bw[3];
    int brute = N < 3000;
    for (int d = 0; d < 3 && !brute; ++d) {
        double lo, hi;
        if (s->per[d]) { lo = s->lo[d]; hi = s->hi[d]; }
        else {
            lo = 1e300; hi = -1e300;
            for (int i = 0; i < N; ++i) { double x = pos[3 * i + d]; if (x < lo) lo = x; if (x > hi) hi = x; }
            hi += 1e-9 * (fabs(hi) + 1.0);
        }
        int k = (int)floor((hi - lo) / s->h);
        if (k < 1) k = 1;
        if (s->per[d] && k < 3) brute = 1;
        if (k > 2048) k = 2048;
        nb[d] = k; blo[d] = lo; bw[d] = (hi - lo) / k;
    }
    int* cnt = (int*)calloc((size_t)N, sizeof(int));
    if (brute) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < N; ++i) {
            int c = 0;
            for (int j = 0; j < N; ++j) {
                if (j == i) continue;
                double r[3]; rel_vec(s, pos + 3 * i, pos + 3 * j, r);
                if (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < h2) ++c;
            }
            cnt[i] = c;
        }
        s->nl_start[0] = 0;
        for (int i = 0; i < N; ++i) s->nl_start[i + 1] = s->nl_start[i] + cnt[i];
        size_t tot = (size_t)s->nl_start[N];
        if (tot > s->nl_cap) { free(s->nl_j); s->nl_cap = tot + tot / 4 + 16; s->nl_j = (int*)malloc(sizeof(int) * s->nl_cap); }
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < N; ++i) {
            int c = s->nl_start[i];
            for (int j = 0; j < N; ++j) {
                if (j == i) continue;
                double r[3]; rel_vec(s, pos + 3 * i, pos + 3 * j, r);
                if (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < h2) s->nl_j[c++] = j;
            }
        }
        free(cnt);
        return;
    }
    int nbin = nb[0] * nb[1] * nb[2];
    int* bin = (int*)malloc(sizeof(int) * (size_t)N);
    int* bstart = (int*)calloc((size_t)nbin + 1, sizeof(int));
    int* border = (int*)malloc(sizeof(int) * (size_t)N);
    for (int i = 0; i < N; ++i) {
        int c[3];
        for (int d = 0; d < 3; ++d) {
            int k = (int)floor((pos[3 * i + d] - blo[d]) / bw[d]);
            if (k < 0) k = 0;
            if (k >= nb[d]) k = nb[d] - 1;
            c[d] = k;
        }
        bin[i] = c[0] + nb[0] * (c[1] + nb[1] * c[2]);
        bstart[bin[i] + 1]++;
    }
    for (int b = 0; b < nbin; ++b) bstart[b + 1] += bstart[b];
    {
        int* fill = (int*)malloc(sizeof(int) * (size_t)nbin);
        memcpy(fill, bstart, sizeof(int) * (size_t)nbin);
        for (int i = 0; i < N; ++i) border[fill[bin[i]]++] = i;
        free(fill);
    }
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            s->nl_start[0] = 0;
            for (int i = 0; i < N; ++i) s->nl_start[i + 1] = s->nl_start[i] + cnt[i];
            size_t tot = (size_t)s->nl_start[N];
            if (tot > s->nl_cap) { free(s->nl_j); s->nl_cap = tot + tot / 4 + 16; s->nl_j = (int*)malloc(sizeof(int) * s->nl_cap); }
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < N; ++i) {
            int b = bin[i];
            int cx = b % nb[0], cy = (b / nb[0]) % nb[1], cz = b / (nb[0] * nb[1]);
            int c = 0, w = pass ? s->nl_start[i] : 0;
            for (int dz = -1; dz <= 1; ++dz) {
                int z = cz + dz;
                if (s->per[2]) z = (z + nb[2]) % nb[2]; else if (z < 0 || z >= nb[2]) continue;
                for (int dy = -1; dy <= 1; ++dy) {
                    int y = cy + dy;
                    if (s->per[1]) y = (y + nb[1]) % nb[1]; else if (y < 0 || y >= nb[1]) continue;
                    for (int dx = -1; dx <= 1; ++dx) {
                        int x = cx + dx;
                        if (s->per[0]) x = (x + nb[0]) % nb[0]; else if (x < 0 || x >= nb[0]) continue;
                        int bb = x + nb[0] * (y + nb[1] * z);
                        for (int q = bstart[bb]; q < bstart[bb + 1]; ++q) {
                            int j = border[q];
                            if (j == i) continue;
                            double r[3]; rel_vec(s, pos + 3 * i, pos + 3 * j, r);
                            if (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < h2) {
                                if (pass) s->nl_j[w++] = j; else ++c;
                            }
                        }
                    }
                }
            }
            if (!pass) cnt[i] = c;
        }
    }
    free(bin); free(bstart); free(border); free(cnt);
}

/* ------------------------------------------------------------------ */
/* Integrator (Groot & Warren 1997, JCP 107:4423) mapped onto
 * UpdateFluidCoor / PredictFluidVel / CorrectFliudVel (ParticleGPU.cuh:11,14,21). */
/* Moving Bound2 walls (plane Couette shear, an extension: the reference's
 * walls are frozen) translate along x at their wall velocity. */
static int walls_move(const orc_sys* s) { return s->wall_u_lo != 0.0 || s->wall_u_hi != 0.0; }
static double wall_u(const orc_sys* s, double y)
{
    return (float)y >= (float)s->wall_split ? s->wall_u_hi : s->wall_u_lo;
}

void orc_update_coor(const orc_sys* s, double* pos, const double* vel, const double* force)
{
    const double dt = s->dt, a = 0.5 * dt * dt / s->m;
#pragma omp parallel for
    for (int i = s->nf; i < s->N; ++i)
        for (int d = 0; d < 3; ++d)
            pos[3 * i + d] += dt * vel[3 * i + d] + a * force[3 * i + d];
    if (walls_move(s))
        for (int i = s->nb2; i < s->nf; ++i) pos[3 * i] += dt * wall_u(s, pos[3 * i + 1]);
}

/* AppFluidPeriodicBC (ParticleGPU.cuh:12): one wrap per step and axis. */
void orc_periodic_bc(const orc_sys* s, double* pos, int* img)
{
    const int first = walls_move(s) ? s->nb2 : s->nf;
#pragma omp parallel for
    for (int i = first; i < s->N; ++i)
        for (int d = 0; d < 3; ++d) {
            if (!s->per[d]) continue;
            double x = pos[3 * i + d];
            if (x < s->lo[d]) { pos[3 * i + d] = x + s->L[d]; img[3 * i + d] -= 1; }
            else if (x >= s->hi[d]) { pos[3 * i + d] = x - s->L[d]; img[3 * i + d] += 1; }
        }
}

void orc_predict(const orc_sys* s, double* vel, double* velhalf, const double* force)
{
    const double dt = s->dt, im = 1.0 / s->m;
#pragma omp parallel for
    for (int i = s->nf; i < s->N; ++i)
        for (int d = 0; d < 3; ++d) {
            double v = vel[3 * i + d], f = force[3 * i + d] * im;
            velhalf[3 * i + d] = v + 0.5 * dt * f;
            vel[3 * i + d] = v + s->lam * dt * f;
        }
}

void orc_correct(const orc_sys* s, double* vel, const double* velhalf, const double* force)
{
    const double dt = s->dt, im = 1.0 / s->m;
#pragma omp parallel for
    for (int i = s->nf; i < s->N; ++i)
        for (int d = 0; d < 3; ++d)
            vel[3 * i + d] = velhalf[3 * i + d] + 0.5 * dt * force[3 * i + d] * im;
}

/* ------------------------------------------------------------------ */
/* Density (CalFluidDen, ParticleGPU.cuh:13): number density including the
 * self term over SDPD particles (Bound2, fluid, membrane); Bound1 only
 * repels.  Walls keep their density. */
static void density_nl(orc_sys* s, const double* pos, double* rho)
{
    const double w0 = orc_lucy_w(0.0, s->h);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = s->nf; i < s->N; ++i) {
        double d = w0;
        for (int q = s->nl_start[i]; q < s->nl_start[i + 1]; ++q) {
            int j = s->nl_j[q];
            if (j < s->nb2) continue;
            double r[3]; rel_vec(s, pos + 3 * i, pos + 3 * j, r);
            d += orc_lucy_w(sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]), s->h);
        }
        rho[i] = s->m * d;
    }
}
void orc_density(orc_sys* s, const double* pos, double* rho)
{
    build_nlist(s, pos);
    density_nl(s, pos, rho);
}

/* Bound2 velocity (CalBound2Vel, ParticleGPU.cuh:15): no-slip image
 * v_w = 2 u_w - sum W v_j / sum W over moving neighbours (Adami, Hu & Adams
 * 2012 style), u_w the wall's prescribed x velocity (0: wall at rest). */
static void bound2_vel_nl(orc_sys* s, const double* pos, double* vel)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = s->nb2; i < s->nf; ++i) {
        double sw = 0, sv[3] = {0, 0, 0};
        for (int q = s->nl_start[i]; q < s->nl_start[i + 1]; ++q) {
            int j = s->nl_j[q];
            if (j < s->nf) continue;
            double r[3]; rel_vec(s, pos + 3 * i, pos + 3 * j, r);
            double w = orc_lucy_w(sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]), s->h);
            sw += w;
            for (int d = 0; d < 3; ++d) sv[d] += w * vel[3 * j + d];
        }
        /* no-slip image about the wall's own velocity u_w (x only):
         * 2 u_w - sum(W v) / sum(W); u_w with no moving neighbour */
        const double uw = wall_u(s, pos[3 * i + 1]);
        for (int d = 0; d < 3; ++d) vel[3 * i + d] = sw > 0 ? -sv[d] / sw : 0.0;
        vel[3 * i] += sw > 0 ? 2.0 * uw : uw;
    }
}
void orc_bound2_vel(orc_sys* s, const double* pos, double* vel)
{
    build_nlist(s, pos);
    bound2_vel_nl(s, pos, vel);
}

/* EOS (ParticleGPU.cuh:16): linear P = c^2 rho, c1 for walls and fluid,
 * c2 for membrane vertices. */
void orc_eos(const orc_sys* s, const double* rho, double* press)
{
#pragma omp parallel for
    for (int i = 0; i < s->N; ++i) {
        double c = i < s->nc ? s->c1 : s->c2;
        press[i] = c * c * rho[i];
    }
}

/* SDPD pair force (Espanol & Revenga 2003, PRE 67:026705), number
 * densities d = rho/m, F(r) = -W'/r:
 *   fC = (P_i/d_i^2 + P_j/d_j^2) F r_ij
 *   fD = -F/(d_i d_j) [a v_ij + b e (e.v_ij)],  a = 5eta/3 - zeta, b = 5(zeta + eta/3)
 *   fR = s_ij sqrt(2 kT F/(d_i d_j)/dt) [sqrt(a)(xi - e(e.xi)) + sqrt(a+b) e(e.xi)]
 * with xi the pair's three Philox normals and s_ij = +1 for the lower id,
 * -1 for the higher: the random pair force is antisymmetric and its
 * covariance is 2 kT times the friction matrix (fluctuation-dissipation). */
static void sdpd_force_nl(orc_sys* s, const double* pos, const double* vel,
                          const double* rho, const double* press, uint64_t counter,
                          double* force, double* abs_sum)
{
    const double a = 5.0 * s->eta / 3.0 - s->zeta;
    const double b = 5.0 * (s->zeta + s->eta / 3.0);
    const double sa = a > 0 ? sqrt(a) : 0.0, sab = sqrt(a + b);
    const int noise = s->noise_on && s->kT > 0;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->N; ++i) {
        double f[3] = {0, 0, 0}, as = 0;
        if (i >= s->nf) {
            double di = rho[i] / s->m, Pi = press[i] / (di * di);
            for (int q = s->nl_start[i]; q < s->nl_start[i + 1]; ++q) {
                int j = s->nl_j[q];
                if (j < s->nb2) continue;
                double r[3]; rel_vec(s, pos + 3 * i, pos + 3 * j, r);
                double rr = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                if (rr == 0.0) continue;
                double F = orc_lucy_f(rr, s->h);
                double dj = rho[j] / s->m;
                double cC = (Pi + press[j] / (dj * dj)) * F;
                double e[3] = {r[0] / rr, r[1] / rr, r[2] / rr};
                double v[3] = {vel[3 * i] - vel[3 * j], vel[3 * i + 1] - vel[3 * j + 1],
                               vel[3 * i + 2] - vel[3 * j + 2]};
                double ev = e[0] * v[0] + e[1] * v[1] + e[2] * v[2];
                double g = F / (di * dj);
                double fc[3], fd[3], fr[3] = {0, 0, 0};
                for (int d = 0; d < 3; ++d) {
                    fc[d] = cC * r[d];
                    fd[d] = -g * (a * v[d] + b * e[d] * ev);
                }
                if (noise) {
                    double xi[3];
                    orc_pair_normals(s->seed, counter, (uint32_t)i, (uint32_t)j, xi);
                    double ex = e[0] * xi[0] + e[1] * xi[1] + e[2] * xi[2];
                    double amp = sqrt(2.0 * s->kT * g / s->dt) * (i < j ? 1.0 : -1.0);
                    for (int d = 0; d < 3; ++d)
                        fr[d] = amp * (sa * (xi[d] - e[d] * ex) + sab * e[d] * ex);
                }
                for (int d = 0; d < 3; ++d) f[d] += fc[d] + fd[d] + fr[d];
                as += sqrt(fc[0] * fc[0] + fc[1] * fc[1] + fc[2] * fc[2]) +
                      sqrt(fd[0] * fd[0] + fd[1] * fd[1] + fd[2] * fd[2]) +
                      sqrt(fr[0] * fr[0] + fr[1] * fr[1] + fr[2] * fr[2]);
            }
        }
        force[3 * i] = f[0]; force[3 * i + 1] = f[1]; force[3 * i + 2] = f[2];
        if (abs_sum) abs_sum[i] = as;
    }
}
void orc_sdpd_force(orc_sys* s, const double* pos, const double* vel,
                    const double* rho, const double* press, uint64_t counter,
                    double* force, double* abs_sum)
{
    build_nlist(s, pos);
    sdpd_force_nl(s, pos, vel, rho, press, counter, force, abs_sum);
}

/* Bound1 repulsion (CalFluidRepForce, ParticleGPU.cuh:18; form not in the
 * reference tree): the conservative SDPD term of a pair at rest density
 * under wall pressure P_w = rep_scale c1^2 rho0:
 *   f = 2 P_w / d0^2 F(r) r_ij,  d0 = rho0/m. */
static void rep_force_nl(orc_sys* s, const double* pos, double* force, double* abs_sum)
{
    const double d0 = s->rho0 / s->m;
    const double K = 2.0 * s->rep_scale * s->c1 * s->c1 * s->rho0 / (d0 * d0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = s->nf; i < s->N; ++i) {
        double f[3] = {0, 0, 0}, as = 0;
        for (int q = s->nl_start[i]; q < s->nl_start[i + 1]; ++q) {
            int j = s->nl_j[q];
            if (j >= s->nb2) continue;
            double r[3]; rel_vec(s, pos + 3 * i, pos + 3 * j, r);
            double rr = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            if (rr == 0.0) continue;
            double c = K * orc_lucy_f(rr, s->h);
            for (int d = 0; d < 3; ++d) f[d] += c * r[d];
            as += fabs(c) * rr;
        }
        for (int d = 0; d < 3; ++d) force[3 * i + d] += f[d];
        if (abs_sum) abs_sum[i] += as;
    }
}
void orc_rep_force(orc_sys* s, const double* pos, double* force, double* abs_sum)
{
    build_nlist(s, pos);
    rep_force_nl(s, pos, force, abs_sum);
}

/* Body force (SetFluidExtForce, ParticleGPU.cuh:19) on moving particles. */
void orc_ext_force(const orc_sys* s, const double* pos, int driveon,
                   double* force, double* abs_sum)
{
    if (!driveon) return;
#pragma omp parallel for
    for (int i = s->nf; i < s->N; ++i) {
        double f = s->fext;
        if (s->force_mode == GC_FORCE_REVERSE_POISEUILLE && pos[3 * i + 1] >= s->force_split)
            f = -f;
        force[3 * i] += f;
        if (abs_sum) abs_sum[i] += fabs(f);
    }
}

/* ------------------------------------------------------------------ */
/* Membrane geometry and mechanics.  Vertex v (0-based membrane index) is
 * particle nc + v; unwrapped position = pos + img * L on periodic axes. */
static void unwrap(const orc_sys* s, const double* pos, const int* img, int v, double* u)
{
    int i = s->nc + v;
    for (int d = 0; d < 3; ++d)
        u[d] = pos[3 * i + d] + (s->per[d] && img ? img[3 * i + d] * s->L[d] : 0.0);
}
static void sub3(const double* a, const double* b, double* o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static void cross3(const double* a, const double* b, double* o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double norm3(const double* a) { return sqrt(dot3(a, a)); }

static int cell_of_tri(const orc_sys* s, int t) { return s->celltab[s->tri[3 * t]]; }

/* CalGeoPara (ParticleGPU.cuh:25-26).  Cell volume = sum of signed
 * tetrahedra against the cell's first triangle's first vertex. */
void orc_geo(orc_sys* s, long stepcount, const double* pos, const int* img,
             double* tri_nor, double* tri_area, double* tri_cen,
             double* cell_area, double* cell_vol, double* par_area)
{
    if (s->ncell <= 0) return;
    double* At = (double*)malloc(sizeof(double) * (size_t)s->ntri);
    for (int c = 0; c < s->ncell; ++c) {
        double O[3]; unwrap(s, pos, img, s->tri[3 * s->tb[c]], O);
        double A = 0, V = 0;
        for (int t = s->tb[c]; t <= s->te[c]; ++t) {
            double P0[3], P1[3], P2[3], e1[3], e2[3], n[3];
            unwrap(s, pos, img, s->tri[3 * t], P0);
            unwrap(s, pos, img, s->tri[3 * t + 1], P1);
            unwrap(s, pos, img, s->tri[3 * t + 2], P2);
            sub3(P1, P0, e1); sub3(P2, P0, e2); cross3(e1, e2, n);
            double nn = norm3(n);
            At[t] = 0.5 * nn;
            if (tri_area) tri_area[t] = At[t];
            if (tri_nor) for (int d = 0; d < 3; ++d) tri_nor[3 * t + d] = n[d] / nn;
            if (tri_cen) for (int d = 0; d < 3; ++d) tri_cen[3 * t + d] = (P0[d] + P1[d] + P2[d]) / 3.0;
            double a[3], b[3], c3[3], bc[3];
            sub3(P0, O, a); sub3(P1, O, b); sub3(P2, O, c3); cross3(b, c3, bc);
            A += At[t];
            V += dot3(a, bc) / 6.0;
        }
        s->Ac[c] = A; s->Vc[c] = V;
        if (cell_area) cell_area[c] = A;
        if (cell_vol) cell_vol[c] = V;
        if (stepcount == -1) {
            s->A0c[c] = A; s->V0c[c] = V;
            for (int t = s->tb[c]; t <= s->te[c]; ++t) s->A0t[t] = At[t];
        }
    }
    if (stepcount == -1) s->have_ref = 1;
    if (par_area) {
        for (int v = 0; v < s->nm; ++v) par_area[v] = 0;
        for (int t = 0; t < s->ntri; ++t)
            for (int k = 0; k < 3; ++k) par_area[s->tri[3 * t + k]] += At[t] / 3.0;
    }
    free(At);
}

/* signed dihedral of edge (A,B) with wings C (triangle v) and D (triangle w):
 * n1 = e x (C-A), n2 = (D-A) x e, theta = atan2((n1^ x n2^).e^, n1^.n2^) */
static double dihedral(const double* A, const double* B, const double* C, const double* D,
                       double* gA, double* gB, double* gC, double* gD)
{
    double e[3], ac[3], ad[3], n1[3], n2[3];
    sub3(B, A, e); sub3(C, A, ac); sub3(D, A, ad);
    cross3(e, ac, n1); cross3(ad, e, n2);
    double l = norm3(e), q1 = dot3(n1, n1), q2 = dot3(n2, n2);
    double m1 = sqrt(q1), m2 = sqrt(q2);
    double u1[3] = {n1[0] / m1, n1[1] / m1, n1[2] / m1};
    double u2[3] = {n2[0] / m2, n2[1] / m2, n2[2] / m2};
    double x[3]; cross3(u1, u2, x);
    double th = atan2(dot3(x, e) / l, dot3(u1, u2));
    if (gA) {
        double tC = dot3(ac, e) / (l * l), tD = dot3(ad, e) / (l * l);
        for (int d = 0; d < 3; ++d) {
            gC[d] = -l * n1[d] / q1;
            gD[d] = -l * n2[d] / q2;
            gA[d] = -(1.0 - tC) * gC[d] - (1.0 - tD) * gD[d];
            gB[d] = -tC * gC[d] - tD * gD[d];
        }
    }
    return th;
}

/* SetMechPara (ParticleGPU.cuh:28-29).  Fedosov, Caswell & Karniadakis
 * 2010 (Biophys J 98:2215) WLC-POW with x0 = 1/wlc_ratio:
 *   l_max = wlc_ratio l0,  g = 1/(4(1-x0)^2) - 1/4 + x0,
 *   h = x0/(2(1-x0)^3) - 1/(4(1-x0)^2) + 1/4,
 *   p = kT sqrt(3) (h + (m+1) g) / (4 mu0 l0)   (linear shear modulus mu0),
 *   k_p = (kT/p) g l0^m                         (zero force at l0),
 *   k_b = 2 k_c / sqrt(3),  theta0 = initial dihedral. */
void orc_mech_para(orc_sys* s, const double* pos, const int* img,
                   double* l0, double* lmax, double* kb, double* mexp,
                   double* p, double* kp, double* theta0)
{
    const double x0 = 1.0 / s->wlc_ratio, m = s->pow_index;
    const double g = 1.0 / (4.0 * (1 - x0) * (1 - x0)) - 0.25 + x0;
    const double hh = x0 / (2.0 * (1 - x0) * (1 - x0) * (1 - x0)) - 1.0 / (4.0 * (1 - x0) * (1 - x0)) + 0.25;
    for (int e = 0; e < s->nedge; ++e) {
        const int* E = s->edge + 6 * e;
        double A[3], B[3], C[3], D[3], d[3];
        unwrap(s, pos, img, E[0], A); unwrap(s, pos, img, E[1], B);
        unwrap(s, pos, img, E[2], C); unwrap(s, pos, img, E[3], D);
        sub3(B, A, d);
        double L0 = norm3(d);
        double P = s->kT * sqrt(3.0) * (hh + (m + 1.0) * g) / (4.0 * s->mu0[e] * L0);
        s->l0[e] = L0;
        s->lmax[e] = L0 * s->wlc_ratio;
        s->mexp[e] = m;
        s->p[e] = P;
        s->kp[e] = s->kT / P * g * pow(L0, m);
        s->kb[e] = 2.0 * s->kc[e] / sqrt(3.0);
        s->th0[e] = dihedral(A, B, C, D, NULL, NULL, NULL, NULL);
        if (l0) l0[e] = s->l0[e];
        if (lmax) lmax[e] = s->lmax[e];
        if (kb) kb[e] = s->kb[e];
        if (mexp) mexp[e] = s->mexp[e];
        if (p) p[e] = s->p[e];
        if (kp) kp[e] = s->kp[e];
        if (theta0) theta0[e] = s->th0[e];
    }
    s->have_mech = 1;
}

static void addv(double* f, int v, const double* x, double sgn)
{
    f[3 * v] += sgn * x[0]; f[3 * v + 1] += sgn * x[1]; f[3 * v + 2] += sgn * x[2];
}

/* CalCellDefForce (ParticleGPU.cuh:30-33): F = -grad E with
 *   E = sum_edges [U_wlc(l) + U_pow(l) + k_b (1 - cos(theta - theta0))]
 *     + sum_tri k_d (A_t - A_t0)^2 / (2 A_t0)
 *     + sum_cells [k_a (A - A0)^2 / (2 A0) + k_v (V - V0)^2 / (2 V0)]
 *   U_wlc = kT l_max/(4p) (3x^2 - 2x^3)/(1-x), x = l/l_max
 *   U_pow = k_p / ((m-1) l^(m-1))          (m != 1)
 * Cell totals A, V are those of the last orc_geo call. */
void orc_def_force(orc_sys* s, const double* pos, const int* img,
                   double* defforce, double* force, double* abs_sum)
{
    if (s->ncell <= 0) return;
    const int nm = s->nm;
    double* fv = (double*)calloc((size_t)nm * 3, sizeof(double));
    double* av = (double*)calloc((size_t)nm, sizeof(double));
    for (int e = 0; e < s->nedge; ++e) {
        const int* E = s->edge + 6 * e;
        double A[3], B[3], C[3], D[3], d[3];
        unwrap(s, pos, img, E[0], A); unwrap(s, pos, img, E[1], B);
        unwrap(s, pos, img, E[2], C); unwrap(s, pos, img, E[3], D);
        /* spring on A (B gets the opposite) */
        sub3(A, B, d);
        double l = norm3(d), x = l / s->lmax[e];
        double fm = -(s->kT / s->p[e]) * (0.25 / ((1 - x) * (1 - x)) - 0.25 + x)
                    + s->kp[e] / pow(l, s->mexp[e]);
        double fs[3] = {fm * d[0] / l, fm * d[1] / l, fm * d[2] / l};
        addv(fv, E[0], fs, 1.0); addv(fv, E[1], fs, -1.0);
        av[E[0]] += fabs(fm); av[E[1]] += fabs(fm);
        /* bending */
        double gA[3], gB[3], gC[3], gD[3];
        double th = dihedral(A, B, C, D, gA, gB, gC, gD);
        double cb = -s->kb[e] * sin(th - s->th0[e]);
        addv(fv, E[0], gA, cb); addv(fv, E[1], gB, cb);
        addv(fv, E[2], gC, cb); addv(fv, E[3], gD, cb);
        av[E[0]] += fabs(cb) * norm3(gA); av[E[1]] += fabs(cb) * norm3(gB);
        av[E[2]] += fabs(cb) * norm3(gC); av[E[3]] += fabs(cb) * norm3(gD);
    }
    for (int t = 0; t < s->ntri; ++t) {
        int c = cell_of_tri(s, t);
        int v0 = s->tri[3 * t], v1 = s->tri[3 * t + 1], v2 = s->tri[3 * t + 2];
        double P[3][3], O[3], e1[3], e2[3], n[3];
        unwrap(s, pos, img, v0, P[0]); unwrap(s, pos, img, v1, P[1]); unwrap(s, pos, img, v2, P[2]);
        unwrap(s, pos, img, s->tri[3 * s->tb[c]], O);
        sub3(P[1], P[0], e1); sub3(P[2], P[0], e2); cross3(e1, e2, n);
        double nn = norm3(n), At = 0.5 * nn;
        double un[3] = {n[0] / nn, n[1] / nn, n[2] / nn};
        double cA = s->ka[t] * (s->Ac[c] - s->A0c[c]) / s->A0c[c] + s->kd[t] * (At - s->A0t[t]) / s->A0t[t];
        double cV = s->kv[t] * (s->Vc[c] - s->V0c[c]) / s->V0c[c];
        int vv[3] = {v0, v1, v2};
        for (int k = 0; k < 3; ++k) {
            double dn[3], dA[3], pa[3], pb[3], dV[3];
            sub3(P[(k + 2) % 3], P[(k + 1) % 3], dn);
            cross3(un, dn, dA);
            sub3(P[(k + 1) % 3], O, pa); sub3(P[(k + 2) % 3], O, pb);
            cross3(pa, pb, dV);
            double f[3];
            for (int d2 = 0; d2 < 3; ++d2) f[d2] = -cA * 0.5 * dA[d2] - cV * dV[d2] / 6.0;
            addv(fv, vv[k], f, 1.0);
            av[vv[k]] += norm3(f);
        }
    }
    for (int v = 0; v < nm; ++v) {
        int i = s->nc + v;
        if (defforce) for (int d = 0; d < 3; ++d) defforce[3 * v + d] = fv[3 * v + d];
        if (force) for (int d = 0; d < 3; ++d) force[3 * i + d] += fv[3 * v + d];
        if (abs_sum) abs_sum[i] += av[v];
    }
    free(fv); free(av);
}

double orc_membrane_energy(orc_sys* s, const double* pos, const int* img)
{
    if (s->ncell <= 0) return 0.0;
    double E = 0;
    for (int e = 0; e < s->nedge; ++e) {
        const int* Ed = s->edge + 6 * e;
        double A[3], B[3], C[3], D[3], d[3];
        unwrap(s, pos, img, Ed[0], A); unwrap(s, pos, img, Ed[1], B);
        unwrap(s, pos, img, Ed[2], C); unwrap(s, pos, img, Ed[3], D);
        sub3(A, B, d);
        double l = norm3(d), x = l / s->lmax[e], m = s->mexp[e];
        E += s->kT * s->lmax[e] / (4.0 * s->p[e]) * (3 * x * x - 2 * x * x * x) / (1 - x);
        if (fabs(m - 1.0) > 1e-12) E += s->kp[e] / ((m - 1.0) * pow(l, m - 1.0));
        else E += -s->kp[e] * log(l);
        double th = dihedral(A, B, C, D, NULL, NULL, NULL, NULL);
        E += s->kb[e] * (1.0 - cos(th - s->th0[e]));
    }
    double* Ac = (double*)calloc((size_t)s->ncell, sizeof(double));
    double* Vc = (double*)calloc((size_t)s->ncell, sizeof(double));
    for (int t = 0; t < s->ntri; ++t) {
        int c = cell_of_tri(s, t);
        double P0[3], P1[3], P2[3], O[3], e1[3], e2[3], n[3], a[3], b[3], cc[3], bc[3];
        unwrap(s, pos, img, s->tri[3 * t], P0); unwrap(s, pos, img, s->tri[3 * t + 1], P1);
        unwrap(s, pos, img, s->tri[3 * t + 2], P2); unwrap(s, pos, img, s->tri[3 * s->tb[c]], O);
        sub3(P1, P0, e1); sub3(P2, P0, e2); cross3(e1, e2, n);
        double At = 0.5 * norm3(n);
        E += s->kd[t] * (At - s->A0t[t]) * (At - s->A0t[t]) / (2.0 * s->A0t[t]);
        sub3(P0, O, a); sub3(P1, O, b); sub3(P2, O, cc); cross3(b, cc, bc);
        Ac[c] += At; Vc[c] += dot3(a, bc) / 6.0;
    }
    for (int c = 0; c < s->ncell; ++c) {
        double ka = s->ka[s->tb[c]], kv = s->kv[s->tb[c]];
        E += ka * (Ac[c] - s->A0c[c]) * (Ac[c] - s->A0c[c]) / (2.0 * s->A0c[c]);
        E += kv * (Vc[c] - s->V0c[c]) * (Vc[c] - s->V0c[c]) / (2.0 * s->V0c[c]);
    }
    free(Ac); free(Vc);
    return E;
}

/* Morse aggregation (CalCellAggForce, ParticleGPU.cuh:34-35):
 *   U = De [exp(2 beta (r0 - r)) - 2 exp(beta (r0 - r))],  r < MaxCut,
 * between membrane vertices of different cells; f_i = -dU/dr e_ij. */
static void agg_force_nl(orc_sys* s, const double* pos, double* aggforce,
                         double* force, double* abs_sum)
{
    const double mc2 = s->mcut * s->mcut;
#pragma omp parallel for schedule(dynamic, 64)
    for (int v = 0; v < s->nm; ++v) {
        int i = s->nc + v;
        double f[3] = {0, 0, 0}, as = 0;
        for (int q = s->nl_start[i]; q < s->nl_start[i + 1]; ++q) {
            int j = s->nl_j[q];
            if (j < s->nc) continue;
            if (s->celltab[j - s->nc] == s->celltab[v]) continue;
            double r[3]; rel_vec(s, pos + 3 * i, pos + 3 * j, r);
            double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (r2 >= mc2 || r2 == 0.0) continue;
            double rr = sqrt(r2);
            double ex = exp(s->beta * (s->r0 - rr));
            double fm = 2.0 * s->beta * s->De * (ex * ex - ex);
            for (int d = 0; d < 3; ++d) f[d] += fm * r[d] / rr;
            as += fabs(fm);
        }
        if (aggforce) for (int d = 0; d < 3; ++d) aggforce[3 * v + d] = f[d];
        if (force) for (int d = 0; d < 3; ++d) force[3 * i + d] += f[d];
        if (abs_sum) abs_sum[i] += as;
    }
}
void orc_agg_force(orc_sys* s, const double* pos, double* aggforce, double* force,
                   double* abs_sum)
{
    if (s->ncell <= 0) return;
    build_nlist(s, pos);
    agg_force_nl(s, pos, aggforce, force, abs_sum);
}

double orc_morse_energy(orc_sys* s, const double* pos)
{
    double E = 0;
    for (int v = 0; v < s->nm; ++v)
        for (int w = v + 1; w < s->nm; ++w) {
            if (s->celltab[v] == s->celltab[w]) continue;
            double r[3]; rel_vec(s, pos + 3 * (s->nc + v), pos + 3 * (s->nc + w), r);
            double rr = norm3(r);
            if (rr >= s->mcut) continue;
            double ex = exp(s->beta * (s->r0 - rr));
            E += s->De * (ex * ex - 2.0 * ex);
        }
    return E;
}

/* ------------------------------------------------------------------ */
/* Particle::UpdatePar (Particle.cu:1310-1329). */
void orc_step(orc_sys* s, orc_state* st, int driveon)
{
#ifdef _OPENMP
    if (g_threads > 0) omp_set_num_threads(g_threads);
#endif
    orc_update_coor(s, st->pos, st->vel, st->force);
    orc_periodic_bc(s, st->pos, st->img);
    build_nlist(s, st->pos);
    density_nl(s, st->pos, st->rho);
    orc_predict(s, st->vel, st->velhalf, st->force);
    bound2_vel_nl(s, st->pos, st->vel);
    orc_eos(s, st->rho, st->press);
    sdpd_force_nl(s, st->pos, st->vel, st->rho, st->press, st->counter, st->force, st->abs_sum);
    st->counter += 1;
    rep_force_nl(s, st->pos, st->force, st->abs_sum);
    orc_ext_force(s, st->pos, driveon, st->force, st->abs_sum);
    if (s->ncell > 0) {
        orc_geo(s, st->step, st->pos, st->img, NULL, NULL, NULL, st->cell_area, st->cell_vol, NULL);
        orc_def_force(s, st->pos, st->img, st->defforce, st->force, st->abs_sum);
        agg_force_nl(s, st->pos, st->aggforce, st->force, st->abs_sum);
    }
    orc_correct(s, st->vel, st->velhalf, st->force);
    bound2_vel_nl(s, st->pos, st->vel);
    st->step += 1;
}
