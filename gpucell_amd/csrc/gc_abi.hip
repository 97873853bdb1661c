// gc_abi.hip -- library state (status, stream, parameter block, scratch),
// the scan / counting-sort primitives, and the cell-list entry points that
// replace cudiv::PreConList, Sort, CalBeginEndCube, SortBasicArrays and
// SortArraysWithCells (src/ConfigDomain.cuh:11-15).
#include "gc_kernels.h"

#include <math.h>
#include <stdio.h>
#include <string.h>
#include <mutex>
#include <vector>

namespace gc {

static thread_local std::string g_err = "ok";   // per host thread
static hipStream_t g_stream = nullptr;
static Structconst g_consts;
static bool g_have_consts = false;
static gc_ext_consts g_ext;
static bool g_have_ext = false;

struct ScratchBuf { void* p = nullptr; size_t bytes = 0; };
static ScratchBuf g_scratch[16];

int set_error(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return GC_OK;
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return GC_ERR_HIP;
}

hipStream_t stream() { return g_stream; }

void make_params(const Structconst& c, const gc_ext_consts& ext, GcParams* P)
{
    memset(P, 0, sizeof(*P));
    P->nb2 = (int)c.nBound2Start; P->nf = (int)c.nFluidStart; P->nc = (int)c.nCellStart;
    P->N = (int)c.nTotalPar; P->nm = (int)c.nCellPar;
    const double pi = 3.14159265358979323846;
    double h = c.CutRadius;
    P->h = c.CutRadius; P->h2 = c.CutRadius * c.CutRadius;
    P->m = c.Mass; P->inv_m = 1.0f / c.Mass;
    P->c1sq = c.SoundSpeed1 * c.SoundSpeed1; P->c2sq = c.SoundSpeed2 * c.SoundSpeed2;
    P->rho0 = c.DenInit; P->lam = c.VVsigma; P->fext = c.ConstExtForce; P->dt = c.TimeStep;
    P->wnorm = (float)(105.0 / (16.0 * pi * h * h * h));
    P->fnorm = (float)(315.0 / (4.0 * pi * h * h * h * h * h));
    double a = 5.0 * c.ShearVis / 3.0 - c.BulkVis, b = 5.0 * (c.BulkVis + c.ShearVis / 3.0);
    P->visc_a = (float)a; P->visc_b = (float)b;
    P->sq_a = (float)(a > 0 ? sqrt(a) : 0.0); P->sq_ab = (float)sqrt(a + b > 0 ? a + b : 0.0);
    P->two_kT_over_dt = (float)(2.0 * c.Temp / c.TimeStep);
    double d0 = (double)c.DenInit / c.Mass;
    P->krep = (float)(2.0 * ext.rep_scale * c.SoundSpeed1 * c.SoundSpeed1 * c.DenInit / (d0 * d0));
    P->De = c.surEnergy; P->mcut2 = c.MaxCut * c.MaxCut; P->r0 = c.zeroLen; P->beta = c.scaFac;
    P->kT = c.Temp;
    P->force_split = ext.force_split;
    P->wall_split = ext.wall_split; P->wall_u_lo = ext.wall_u_lo; P->wall_u_hi = ext.wall_u_hi;
    for (int d = 0; d < 3; ++d) {
        P->per[d] = ext.periodic[d];
        P->off[d] = ext.periodic[d] ? 0.0f : c.CutRadius;
    }
    P->force_mode = ext.force_mode;
    P->noise_on = ext.noise_on && c.Temp > 0.0f;
    P->wlc_ratio = ext.wlc_ratio; P->pow_index = ext.pow_index;
    uint64_t seed = (uint64_t)c.iSeed;
    P->seed_lo = (uint32_t)seed; P->seed_hi = (uint32_t)(seed >> 32);
    P->ctr_lo = (uint32_t)ext.noise_counter; P->ctr_hi = (uint32_t)(ext.noise_counter >> 32);
}

int params(GcParams* out)
{
    if (!g_have_consts) return set_error(GC_ERR_NOCONSTS, "gc_ConstsDataUp has not been called");
    if (!g_have_ext) { gc_DefaultExtConsts(&g_ext); g_have_ext = true; }
    make_params(g_consts, g_ext, out);
    return GC_OK;
}

void advance_noise_counter()
{
    if (!g_have_ext) { gc_DefaultExtConsts(&g_ext); g_have_ext = true; }
    g_ext.noise_counter += 1;
}

int scratch(int slot, size_t bytes, void** ptr)
{
    ScratchBuf& b = g_scratch[slot];
    if (bytes == 0) bytes = 256;
    if (b.bytes < bytes) {
        if (b.p) {
            (void)hipStreamSynchronize(g_stream);
            (void)hipFree(b.p);
        }
        b.p = nullptr; b.bytes = 0;
        size_t want = bytes + bytes / 4;
        if (hipMalloc(&b.p, want) != hipSuccess) {
            b.p = nullptr;
            return set_error(GC_ERR_NOMEM, "scratch allocation failed");
        }
        b.bytes = want;
    }
    *ptr = b.p;
    return GC_OK;
}

// ---- scan -------------------------------------------------------------
constexpr int SCAN_BS = 256;
constexpr int SCAN_IPT = 4;
constexpr int SCAN_TILE = SCAN_BS * SCAN_IPT;

// Exclusive scan of one tile per block; block totals to sums[blockIdx].
__global__ __launch_bounds__(SCAN_BS) void k_scan_tile(const int* __restrict__ in,
                                                       int* __restrict__ out,
                                                       int* __restrict__ sums, int n)
{
    __shared__ int lds[SCAN_BS];
    int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_IPT;
    int v[SCAN_IPT];
    int t = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        t += v[k];
    }
    lds[threadIdx.x] = t;
    __syncthreads();
    for (int o = 1; o < SCAN_BS; o <<= 1) {
        int a = threadIdx.x >= o ? lds[threadIdx.x - o] : 0;
        __syncthreads();
        lds[threadIdx.x] += a;
        __syncthreads();
    }
    int run = lds[threadIdx.x] - t;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == SCAN_BS - 1 && sums) sums[blockIdx.x] = lds[SCAN_BS - 1];
}

__global__ void k_scan_add(int* __restrict__ out, const int* __restrict__ offs, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += offs[i / SCAN_TILE];
}

__global__ void k_write_total(int* out, const int* in, int n)
{
    // out[n] = out[n-1] + in[n-1]; one thread
    if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = n > 0 ? out[n - 1] + in[n - 1] : 0;
}

// Writes exclusive prefix sums of in[0..n) to out[0..n) and the total to
// out[n].  `in` and `out` may be the same array only if a copy of the last
// element survives: we read in[n-1] before the final write, so callers pass
// distinct arrays.
static int scan_rec(const int* in, int* out, int n, hipStream_t s, int depth, int* ws)
{
    int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb <= 1) {
        k_scan_tile<<<1, SCAN_BS, 0, s>>>(in, out, nullptr, n);
        GC_LAUNCH_CHECK("scan tile");
        return GC_OK;
    }
    int* sums;
    if (ws) {
        sums = ws;
        ws += 2 * nb + 2;
    } else {
        void* p;
        int st = scratch(8 + depth, sizeof(int) * (size_t)(2 * nb + 2), &p);
        if (st) return st;
        sums = (int*)p;
    }
    int* soffs = sums + nb + 1;
    k_scan_tile<<<nb, SCAN_BS, 0, s>>>(in, out, sums, n);
    GC_LAUNCH_CHECK("scan tile");
    int st = scan_rec(sums, soffs, nb, s, depth + 1, ws);
    if (st) return st;
    k_scan_add<<<blocks_for(n, 256), 256, 0, s>>>(out, soffs, n);
    GC_LAUNCH_CHECK("scan add");
    return GC_OK;
}

size_t scan_ws_ints(int n)
{
    size_t total = 0;
    for (;;) {
        int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
        if (nb <= 1) break;
        total += 2 * (size_t)nb + 2;
        n = nb;
    }
    return total + 1;
}

int exclusive_scan(const int* in, int* out, int n, hipStream_t s, int* ws)
{
    int st = scan_rec(in, out, n, s, 0, ws);
    if (st) return st;
    k_write_total<<<1, 1, 0, s>>>(out, in, n);
    GC_LAUNCH_CHECK("scan total");
    return GC_OK;
}

// ---- counting sort ------------------------------------------------------
__global__ void k_fill_int(int* __restrict__ a, int v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_hist(int n, const int* __restrict__ keys, int* __restrict__ cnt,
                       int* __restrict__ rank)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) rank[p] = atomicAdd(&cnt[keys[p]], 1);
}

__global__ void k_scatter(int n, const int* __restrict__ keys, const int* __restrict__ vals,
                          const int* __restrict__ offs, const int* __restrict__ rank,
                          int* __restrict__ tk, int* __restrict__ tv)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int k = keys[p];
    int q = offs[k] + rank[p];
    tk[q] = k;
    tv[q] = vals[p];
}

// Final place of element q inside its key segment = number of elements
// with a smaller value (ties by position), which makes the order
// independent of the atomic arrival order of k_hist.
__global__ void k_rankfix(int n, const int* __restrict__ tk, const int* __restrict__ tv,
                          const int* __restrict__ offs, int* __restrict__ ok,
                          int* __restrict__ ov)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int k = tk[q], v = tv[q];
    int b = offs[k], e = offs[k + 1];
    int r = 0;
    for (int t = b; t < e; ++t) {
        int w = tv[t];
        r += (w < v) | ((w == v) & (t < q));
    }
    ok[b + r] = k;
    ov[b + r] = v;
}

int counting_sort(int n, int K, const int* keys, const int* vals, int* out_k, int* out_v,
                  int* cnt, int* offs, int* rank, int* tmp_k, int* tmp_v, hipStream_t s,
                  int* scan_ws)
{
    k_fill_int<<<blocks_for(K + 1, 256), 256, 0, s>>>(cnt, 0, K + 1);
    GC_LAUNCH_CHECK("sort fill");
    if (n > 0) {
        k_hist<<<blocks_for(n, 256), 256, 0, s>>>(n, keys, cnt, rank);
        GC_LAUNCH_CHECK("sort hist");
    }
    int st = exclusive_scan(cnt, offs, K, s, scan_ws);
    if (st) return st;
    if (n > 0) {
        k_scatter<<<blocks_for(n, 256), 256, 0, s>>>(n, keys, vals, offs, rank, tmp_k, tmp_v);
        GC_LAUNCH_CHECK("sort scatter");
        k_rankfix<<<blocks_for(n, 256), 256, 0, s>>>(n, tmp_k, tmp_v, offs, out_k, out_v);
        GC_LAUNCH_CHECK("sort rankfix");
    }
    return GC_OK;
}

}  // namespace gc

using namespace gc;

// ---------------------------------------------------------------------------
// Library state entry points
extern "C" const char* gc_GetLastError(void) { return g_err.c_str(); }
extern "C" const char* gc_Version(void) { return "gpucell_amd 0.1 (gfx950)"; }

extern "C" int gc_SetStream(void* s)
{
    g_stream = (hipStream_t)s;
    return GC_OK;
}

extern "C" int gc_Synchronize(void)
{
    return hip_check(hipStreamSynchronize(g_stream), "hipStreamSynchronize");
}

extern "C" int gc_ConstsDataUp(const Structconst* c)
{
    if (!c) return set_error(GC_ERR_INVALID, "ConstsDataUp: null block");
    if (!(c->CutRadius > 0) || !(c->Mass > 0) || !(c->TimeStep > 0) || !(c->DenInit > 0))
        return set_error(GC_ERR_INVALID, "ConstsDataUp: CutRadius, Mass, TimeStep and DenInit must be > 0");
    if (5.0 * c->ShearVis / 3.0 - c->BulkVis < 0)
        return set_error(GC_ERR_INVALID, "ConstsDataUp: need 5*ShearVis/3 >= BulkVis");
    if (c->surEnergy != 0 && c->MaxCut > c->CutRadius)
        return set_error(GC_ERR_GRID, "ConstsDataUp: MaxCut must not exceed CutRadius");
    if (!(c->nBound2Start <= c->nFluidStart && c->nFluidStart <= c->nCellStart &&
          c->nCellStart <= c->nTotalPar && c->nTotalPar - c->nCellStart == c->nCellPar))
        return set_error(GC_ERR_INVALID, "ConstsDataUp: inconsistent id ranges");
    g_consts = *c;
    g_have_consts = true;
    return GC_OK;
}

extern "C" void gc_DefaultExtConsts(gc_ext_consts* e)
{
    memset(e, 0, sizeof(*e));
    e->periodic[0] = 1; e->periodic[1] = 0; e->periodic[2] = 0;
    e->force_mode = GC_FORCE_UNIFORM;
    e->force_split = 0.0f;
    e->noise_on = 1;
    e->wlc_ratio = 2.2f;
    e->pow_index = 2.0f;
    e->rep_scale = 1.0f;
    e->noise_counter = 0;
}

extern "C" int gc_SetExtConsts(const gc_ext_consts* e)
{
    if (!e) return set_error(GC_ERR_INVALID, "SetExtConsts: null");
    if (!(e->wlc_ratio > 1.0f)) return set_error(GC_ERR_INVALID, "SetExtConsts: wlc_ratio must be > 1");
    g_ext = *e;
    g_have_ext = true;
    return GC_OK;
}

extern "C" int gc_GetExtConsts(gc_ext_consts* e)
{
    if (!e) return set_error(GC_ERR_INVALID, "GetExtConsts: null");
    if (!g_have_ext) { gc_DefaultExtConsts(&g_ext); g_have_ext = true; }
    *e = g_ext;
    return GC_OK;
}

extern "C" int gc_ConfigDomain(const float bbox[6], float cut, const int periodic[3],
                               int ncube[3], float outer[3], float invwid[3], int* sizecube)
{
    if (!bbox || !periodic || !ncube || !outer || !invwid || !sizecube || !(cut > 0))
        return set_error(GC_ERR_INVALID, "ConfigDomain: bad arguments");
    long K = 1;
    for (int d = 0; d < 3; ++d) {
        float L = bbox[2 * d + 1] - bbox[2 * d];
        if (!periodic[d]) L = L + 2.0f * cut;
        outer[d] = L;
        ncube[d] = (int)floorf(L / cut);
        if (ncube[d] < 1) return set_error(GC_ERR_GRID, "ConfigDomain: box smaller than one cube");
        if (periodic[d] && ncube[d] < 3)
            return set_error(GC_ERR_GRID, "ConfigDomain: a periodic axis needs >= 3 cubes");
        invwid[d] = (float)ncube[d] / outer[d];
        K *= ncube[d];
    }
    if (K > (1L << 30)) return set_error(GC_ERR_GRID, "ConfigDomain: too many cubes");
    *sizecube = (int)K;
    return GC_OK;
}

// ---------------------------------------------------------------------------
// cudiv::PreConList
__global__ void k_preconlist(int n, GcParams P, const tfloat3* __restrict__ coor,
                             const float* __restrict__ bbox, const float* __restrict__ invw,
                             const int* __restrict__ ncube, tint3* __restrict__ cc,
                             int* __restrict__ cubepart, const int* __restrict__ sortpart,
                             int* __restrict__ oldpos)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    tfloat3 x = coor[p];
    int nx = ncube[0], ny = ncube[1], nz = ncube[2];
    int cx = cube_coord(x.x, bbox[0], P.off[0], invw[0], nx);
    int cy = cube_coord(x.y, bbox[2], P.off[1], invw[1], ny);
    int cz = cube_coord(x.z, bbox[4], P.off[2], invw[2], nz);
    cc[p] = tint3{cx, cy, cz};
    cubepart[p] = cx + nx * (cy + ny * cz);
    oldpos[sortpart[p]] = p;
}

extern "C" int gc_PreConList(int n, int cutradius, tfloat3* d_totalcoor, float* d_bboxsph,
                             float* d_invwid, int* d_ncube, tint3* d_cc, int* d_cubepart,
                             int* d_sortpart, int* d_oldpos)
{
    (void)cutradius;
    GcParams P;
    int st = params(&P);
    if (st) return st;
    if (n < 0) return set_error(GC_ERR_INVALID, "PreConList: n < 0");
    if (n == 0) return GC_OK;
    if (!d_totalcoor || !d_bboxsph || !d_invwid || !d_ncube || !d_cc || !d_cubepart ||
        !d_sortpart || !d_oldpos)
        return set_error(GC_ERR_INVALID, "PreConList: null pointer");
    k_preconlist<<<blocks_for(n, 256), 256, 0, stream()>>>(n, P, d_totalcoor, d_bboxsph,
        d_invwid, d_ncube, d_cc, d_cubepart, d_sortpart, d_oldpos);
    GC_LAUNCH_CHECK("PreConList");
    return GC_OK;
}

// ---------------------------------------------------------------------------
// cudiv::Sort
__global__ void k_max_int(int n, const int* __restrict__ a, int* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int v = i < n ? a[i] : -1;
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0 && v >= 0) atomicMax(out, v);
}

__global__ void k_min_int(int n, const int* __restrict__ a, int* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int v = i < n ? a[i] : 0x7fffffff;
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) atomicMin(out, v);
}

// Largest cube population the rank-fix pass accepts (its cost is the
// square of the population).
static const int GC_MAX_CUBE_POP = 8192;

// Buffers of a sort of n elements with keys in [0, K): red holds 4 ints,
// cnt and offs K + 1, rank, tk and tv n each, ws scan_ws_ints(K) (NULL:
// library scratch for the scan).
struct SortBufs {
    int *red = nullptr, *cnt = nullptr, *offs = nullptr, *rank = nullptr, *tk = nullptr, *tv = nullptr;
    int* ws = nullptr;
};

// Smallest and largest key (one sync); K = largest + 1.
static int sort_key_range(int n, const int* keys, int* red, hipStream_t s, int* K)
{
    int init[2] = {-1, 0x7fffffff};
    int st = hip_check(hipMemcpyAsync(red, init, sizeof(init), hipMemcpyHostToDevice, s), "Sort init");
    if (st) return st;
    k_max_int<<<blocks_for(n, 256), 256, 0, s>>>(n, keys, red);
    k_min_int<<<blocks_for(n, 256), 256, 0, s>>>(n, keys, red + 1);
    GC_LAUNCH_CHECK("Sort minmax");
    int mm[2];
    st = hip_check(hipMemcpyAsync(mm, red, sizeof(mm), hipMemcpyDeviceToHost, s), "Sort D2H");
    if (st) return st;
    st = hip_check(hipStreamSynchronize(s), "Sort sync");
    if (st) return st;
    if (mm[1] < 0) return set_error(GC_ERR_INVALID, "Sort: negative key");
    *K = mm[0] + 1;
    return GC_OK;
}

// Population check (largest key count, one sync), then the counting sort
// by (key, value) in place.
static int sort_run(int n, int K, int* keys, int* values, const SortBufs& b, hipStream_t s)
{
    k_fill_int<<<blocks_for(K + 1, 256), 256, 0, s>>>(b.cnt, 0, K + 1);
    k_hist<<<blocks_for(n, 256), 256, 0, s>>>(n, keys, b.cnt, b.rank);
    int neg1 = -1;
    int st = hip_check(hipMemcpyAsync(b.red, &neg1, sizeof(int), hipMemcpyHostToDevice, s), "Sort init2");
    if (st) return st;
    k_max_int<<<blocks_for(K, 256), 256, 0, s>>>(K, b.cnt, b.red);
    GC_LAUNCH_CHECK("Sort population");
    int maxpop = 0;
    st = hip_check(hipMemcpyAsync(&maxpop, b.red, sizeof(int), hipMemcpyDeviceToHost, s), "Sort D2H2");
    if (st) return st;
    st = hip_check(hipStreamSynchronize(s), "Sort sync2");
    if (st) return st;
    if (maxpop > GC_MAX_CUBE_POP)
        return set_error(GC_ERR_GRID, "Sort: a key holds more than 8192 elements (over-full cube)");
    return counting_sort(n, K, keys, values, keys, values, b.cnt, b.offs, b.rank, b.tk, b.tv, s, b.ws);
}

extern "C" int gc_Sort(int n, int* keys, int* values)
{
    if (n < 0) return set_error(GC_ERR_INVALID, "Sort: n < 0");
    if (n == 0) return GC_OK;
    if (!keys || !values) return set_error(GC_ERR_INVALID, "Sort: null pointer");
    hipStream_t s = stream();
    SortBufs b;
    void* p;
    int st = scratch(0, sizeof(int) * 4, &p);
    if (st) return st;
    b.red = (int*)p;
    int K = 0;
    if ((st = sort_key_range(n, keys, b.red, s, &K))) return st;
    void *pc, *po, *pr, *ptk, *ptv;
    if ((st = scratch(1, sizeof(int) * (size_t)(K + 1), &pc))) return st;
    if ((st = scratch(2, sizeof(int) * (size_t)(K + 1), &po))) return st;
    if ((st = scratch(3, sizeof(int) * (size_t)n, &pr))) return st;
    if ((st = scratch(4, sizeof(int) * (size_t)n, &ptk))) return st;
    if ((st = scratch(5, sizeof(int) * (size_t)n, &ptv))) return st;
    b.cnt = (int*)pc; b.offs = (int*)po; b.rank = (int*)pr; b.tk = (int*)ptk; b.tv = (int*)ptv;
    return sort_run(n, K, keys, values, b, s);
}

// ---------------------------------------------------------------------------
// gc_SortNoAlloc: gc_Sort with every buffer in one fixed device array.
// Sizes: n <= GC_NOALLOC_MAX_N elements, keys < GC_NOALLOC_MAX_K.  The array
// holds the 4 range/population words, the key counts and offsets (K + 1
// each), the ranks and the two scatter arrays (n each) and the scan
// workspace, laid out from its start for each call.
namespace noalloc {
constexpr int SCAN_WS = 16384;   // >= scan_ws_ints(GC_NOALLOC_MAX_K) = 8205 (checked per call)
constexpr size_t INTS = 4 + 2 * ((size_t)GC_NOALLOC_MAX_K + 1) + 3 * (size_t)GC_NOALLOC_MAX_N + SCAN_WS;
// One copy per GPU, for the life of the process (84 MB).  In host code this
// name is a stand-in, not a device address: the address is taken once per
// device with hipGetSymbolAddress (noalloc_base) and only that is passed on.
static __device__ __align__(256) int scratch[INTS];
}  // namespace noalloc

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
static void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The current device's address of `scratch`, looked up on the first call
// on that device.
int noalloc_base(int** base)
{
    int dev = 0;
    int st = hip_check(hipGetDevice(&dev), "SortNoAlloc: hipGetDevice");
    if (st) return st;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return set_error(GC_ERR_INVALID, "SortNoAlloc: device index");
    if (!g_noalloc_ptr[dev]) {
        void* ptr = nullptr;
        st = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc::scratch)), "SortNoAlloc: hipGetSymbolAddress");
        if (st) return st;
        if (!ptr) return set_error(GC_ERR_HIP, "SortNoAlloc: no device address for the scratch array");
        g_noalloc_ptr[dev] = ptr;
    }
    *base = (int*)g_noalloc_ptr[dev];
    return GC_OK;
}
}  // namespace

extern "C" int gc_SortNoAlloc(int n, int* keys, int* values)
{
    if (n < 0) return set_error(GC_ERR_INVALID, "SortNoAlloc: n < 0");
    if (n == 0) return GC_OK;
    if (!keys || !values) return set_error(GC_ERR_INVALID, "SortNoAlloc: null pointer");
    if (n > GC_NOALLOC_MAX_N)
        return set_error(GC_ERR_INVALID, "SortNoAlloc: more than GC_NOALLOC_MAX_N elements");
    hipStream_t s = stream();
    int* base = nullptr;
    int st = noalloc_base(&base);
    if (st) return st;
    SortBufs b;
    b.red = base;
    int K = 0;
    if ((st = sort_key_range(n, keys, b.red, s, &K))) return st;
    if (K > GC_NOALLOC_MAX_K)
        return set_error(GC_ERR_INVALID, "SortNoAlloc: a key >= GC_NOALLOC_MAX_K");
    if (scan_ws_ints(K) > (size_t)noalloc::SCAN_WS)
        return set_error(GC_ERR_INVALID, "SortNoAlloc: scan workspace too small");
    int* q = base + 4;
    b.cnt = q; q += K + 1;
    b.offs = q; q += K + 1;
    b.rank = q; q += n;
    b.tk = q; q += n;
    b.tv = q; q += n;
    b.ws = q;
    return sort_run(n, K, keys, values, b, s);
}

// ---------------------------------------------------------------------------
// cudiv::CalBeginEndCube: binary search per cube in the sorted keys.
__global__ void k_begin_end(int n, int K, const int* __restrict__ keys, tint2* __restrict__ be)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    int lo = 0, hi = n;
    while (lo < hi) { int m = (lo + hi) >> 1; if (keys[m] < k) lo = m + 1; else hi = m; }
    int b = lo;
    hi = n;
    while (lo < hi) { int m = (lo + hi) >> 1; if (keys[m] <= k) lo = m + 1; else hi = m; }
    be[k] = tint2{b, lo};
}

extern "C" int gc_CalBeginEndCube(int n, int sizecube, int* d_cubepart, tint2* d_beginendcube)
{
    if (n < 0 || sizecube < 0) return set_error(GC_ERR_INVALID, "CalBeginEndCube: negative size");
    if (sizecube == 0) return GC_OK;
    if (!d_beginendcube || (n > 0 && !d_cubepart))
        return set_error(GC_ERR_INVALID, "CalBeginEndCube: null pointer");
    k_begin_end<<<blocks_for(sizecube, 256), 256, 0, stream()>>>(n, sizecube, d_cubepart,
                                                                  d_beginendcube);
    GC_LAUNCH_CHECK("CalBeginEndCube");
    return GC_OK;
}

// ---------------------------------------------------------------------------
// cudiv::SortBasicArrays / SortArraysWithCells: gather into scratch, copy back.
struct GatherArgs {
    tfloat3* coor; tint3* cc; tfloat4* velrhop; tfloat3* velhalf; tfloat3* force;
    float* press; tint3* nthro; int* curpos;
};

__global__ void k_gather_arrays(int n, const int* __restrict__ sortpart,
                                const int* __restrict__ oldpos, GatherArgs a,
                                GatherArgs t)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int id = sortpart[s];
    int o = oldpos[id];
    t.coor[s] = a.coor[o];
    t.cc[s] = a.cc[o];
    t.velrhop[s] = a.velrhop[o];
    t.velhalf[s] = a.velhalf[o];
    t.force[s] = a.force[o];
    t.press[s] = a.press[o];
    if (a.nthro) t.nthro[s] = a.nthro[o];
    if (a.curpos) a.curpos[id] = s;
}

static int sort_arrays(int n, int* sortpart, int* oldpos, int* curpos, tfloat3* coor,
                       tint3* cc, tfloat4* velrhop, tfloat3* velhalf, tfloat3* force,
                       float* press, tint3* nthro)
{
    if (n < 0) return set_error(GC_ERR_INVALID, "SortArrays: n < 0");
    if (n == 0) return GC_OK;
    if (!sortpart || !oldpos || !coor || !cc || !velrhop || !velhalf || !force || !press)
        return set_error(GC_ERR_INVALID, "SortArrays: null pointer");
    hipStream_t s = stream();
    size_t per = sizeof(tfloat3) * 3 + sizeof(tint3) * 2 + sizeof(tfloat4) + sizeof(float);
    void* p;
    int st = scratch(6, per * (size_t)n + 1024, &p);
    if (st) return st;
    char* b = (char*)p;
    GatherArgs a{coor, cc, velrhop, velhalf, force, press, nthro, curpos};
    GatherArgs t;
    t.velrhop = (tfloat4*)b; b += sizeof(tfloat4) * (size_t)n;
    t.coor = (tfloat3*)b; b += sizeof(tfloat3) * (size_t)n;
    t.velhalf = (tfloat3*)b; b += sizeof(tfloat3) * (size_t)n;
    t.force = (tfloat3*)b; b += sizeof(tfloat3) * (size_t)n;
    t.cc = (tint3*)b; b += sizeof(tint3) * (size_t)n;
    t.nthro = (tint3*)b; b += sizeof(tint3) * (size_t)n;
    t.press = (float*)b;
    t.curpos = nullptr;
    k_gather_arrays<<<blocks_for(n, 256), 256, 0, s>>>(n, sortpart, oldpos, a, t);
    GC_LAUNCH_CHECK("SortArrays gather");
    const hipMemcpyKind dd = hipMemcpyDeviceToDevice;
    st = hip_check(hipMemcpyAsync(coor, t.coor, sizeof(tfloat3) * n, dd, s), "copy coor");
    if (!st) st = hip_check(hipMemcpyAsync(cc, t.cc, sizeof(tint3) * n, dd, s), "copy cc");
    if (!st) st = hip_check(hipMemcpyAsync(velrhop, t.velrhop, sizeof(tfloat4) * n, dd, s), "copy velrhop");
    if (!st) st = hip_check(hipMemcpyAsync(velhalf, t.velhalf, sizeof(tfloat3) * n, dd, s), "copy velhalf");
    if (!st) st = hip_check(hipMemcpyAsync(force, t.force, sizeof(tfloat3) * n, dd, s), "copy force");
    if (!st) st = hip_check(hipMemcpyAsync(press, t.press, sizeof(float) * n, dd, s), "copy press");
    if (!st && nthro) st = hip_check(hipMemcpyAsync(nthro, t.nthro, sizeof(tint3) * n, dd, s), "copy nthro");
    return st;
}

extern "C" int gc_SortBasicArrays(int n, int* d_sortpart, int* d_oldpos, tfloat3* d_totalcoor,
                                  tint3* d_cc, tfloat4* d_velrhop, tfloat3* d_velhalf,
                                  tfloat3* d_totalforce, float* d_press)
{
    return sort_arrays(n, d_sortpart, d_oldpos, nullptr, d_totalcoor, d_cc, d_velrhop,
                       d_velhalf, d_totalforce, d_press, nullptr);
}

extern "C" int gc_SortArraysWithCells(int n, int* d_sortpart, int* d_oldpos, int* d_curpos,
                                      tfloat3* d_totalcoor, tint3* d_cc, tfloat4* d_velrhop,
                                      tfloat3* d_velhalf, tfloat3* d_totalforce,
                                      float* d_press, tint3* d_nthrobound)
{
    if (!d_curpos || !d_nthrobound)
        return set_error(GC_ERR_INVALID, "SortArraysWithCells: null curpos/nthrobound");
    return sort_arrays(n, d_sortpart, d_oldpos, d_curpos, d_totalcoor, d_cc, d_velrhop,
                       d_velhalf, d_totalforce, d_press, d_nthrobound);
}
