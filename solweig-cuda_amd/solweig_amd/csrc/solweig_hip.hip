// MI355X (gfx950) kernels of the SOLWEIG 2022a per-pixel-hour radiation path and their C ABI.
//
// Layout: every raster is a row-major plane [rows][cols]; a domain is described by an
// sw_domain record in device memory; per-(domain, hour) scalars live in an FP64 table of
// SW_HOUR_STRIDE values built on the host. Work is tiled W columns x 256 / W rows per 256-thread
// workgroup (W in 16..128, chosen per domain at run time), tiles of all domains of a launch are
// concatenated, and the workgroup index is remapped so that consecutive tiles stay on one XCD
// (one L2) of the 8. The standalone sub-model kernels at the end map one thread to one pixel.
//
// Exactness: the building/vegetation march (shadows and sky-view visibility) compares FP64
// values formed exactly as the NumPy reference forms them (float32 rasters widened to FP64,
// vegetation DSMs as FP64 sums, no contraction), so the masks are bit-identical to
// umep_solweig.py. Radiation sums run in FP32; wedge, lag blend, surface temperature and the
// UTCI polynomial evaluation run in FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "../../../include/solweig_hip.h"

static_assert(sizeof(sw_domain) == 336, "sw_domain must stay 42 eight-byte fields");

namespace {

constexpr int TILE_X = 32;
constexpr int TILE_Y = 8;
constexpr int BLOCK = TILE_X * TILE_Y;
constexpr int NXCD = 8;
constexpr double SBC = 5.67051e-8;
constexpr int MAX_SECOND = 64;

thread_local std::string g_error;

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

int hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return SW_OK;
    return fail(SW_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Workgroup i is dispatched to XCD i % 8: give each XCD a contiguous run of tiles.
__device__ __forceinline__ int xcd_remap(int i, int total) {
    const int q = total / NXCD, r = total % NXCD;
    const int x = i % NXCD, k = i / NXCD;
    return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

__device__ __forceinline__ int find_domain(const int32_t *bs, int nd, int b) {
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bs[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

struct PixelAt {
    int d, r, c;
    long long p;
    bool valid;
    int tid;                    // the lane's index in the workgroup (threadIdx.x)
    int wr0, wr1, wc0, wc1;     // rows and columns spanned by this pixel's wave (uniform)
    int tr0, tc0, tx, ty;       // the workgroup's tile: first row and column, width, height
};

// Address spaces: per-pixel planes are read and written through global (1) pointers; every
// value that is the same for the whole launch (the domain record, hour record, ray, patch and
// ground-view tables) is read through constant (4) pointers so it lands in scalar registers.
#define AS1 __attribute__((address_space(1)))
#define AS4 __attribute__((address_space(4)))
#define AS3 __attribute__((address_space(3)))
typedef float v2f __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));

struct DomainG {
    int64_t rows, cols, pixels, landcover;
    double dsm_raise;
    int tile_shift;                 // log2 of the tile width
    const AS1 float *dsm, *trees, *walls, *aspect, *emis, *alb, *tgk, *tstart, *tmaxlst, *svfbuveg;
    const AS1 double *tan_asvf, *tan_svfalfa;
    const AS1 uint8_t *ground;
    const AS1 uint32_t *vis;
    AS1 v2f *rec;
    AS1 float *lag, *tmrt, *utci;
    const AS4 double *hours;
    const AS4 float *ptab_f;
    const AS4 double *ptab_d;
    const AS4 int32_t *ray_off;
    const AS4 double *ray_dz;
    AS1 uint8_t *masks;
    AS1 double *shadow_planes;
    AS1 float *full;
    const AS1 float *skysums;
    const AS4 float *hmax;
    int hmax_cols;
    const AS1 uint32_t *gvf_len;
    const AS1 float *gvf_nosh;
    const AS1 float *march_trees;
    const AS1 uint32_t *ptab_s;
    const AS1 uint32_t *gvf_edge;
    const AS1 int8_t *diffuse_tab;
    AS1 v2f *diffuse;
};

__device__ __forceinline__ DomainG domain_at(const sw_domain *D) {
    const AS4 sw_domain *d = (const AS4 sw_domain *)D;
    DomainG g;
    g.rows = d->rows; g.cols = d->cols; g.pixels = d->pixels; g.landcover = d->landcover;
    g.dsm_raise = d->dsm_raise;
    const int64_t tc = d->tile_cols;
    g.tile_shift = tc == 16 ? 4 : (tc == 64 ? 6 : (tc == 128 ? 7 : 5));
    g.dsm = (const AS1 float *)d->dsm; g.trees = (const AS1 float *)d->trees;
    g.walls = (const AS1 float *)d->walls; g.aspect = (const AS1 float *)d->aspect;
    g.emis = (const AS1 float *)d->emis; g.alb = (const AS1 float *)d->alb;
    g.tgk = (const AS1 float *)d->tgk; g.tstart = (const AS1 float *)d->tstart;
    g.tmaxlst = (const AS1 float *)d->tmaxlst; g.svfbuveg = (const AS1 float *)d->svfbuveg;
    g.tan_asvf = (const AS1 double *)d->tan_asvf; g.tan_svfalfa = (const AS1 double *)d->tan_svfalfa;
    g.ground = (const AS1 uint8_t *)d->ground; g.vis = (const AS1 uint32_t *)d->vis;
    g.rec = (AS1 v2f *)d->rec; g.lag = (AS1 float *)d->lag;
    g.tmrt = (AS1 float *)d->tmrt; g.utci = (AS1 float *)d->utci;
    g.hours = (const AS4 double *)d->hours;
    g.ptab_f = (const AS4 float *)d->ptab_f; g.ptab_d = (const AS4 double *)d->ptab_d;
    g.ray_off = (const AS4 int32_t *)d->ray_off; g.ray_dz = (const AS4 double *)d->ray_dz;
    g.masks = (AS1 uint8_t *)d->masks; g.shadow_planes = (AS1 double *)d->shadow_planes;
    g.full = (AS1 float *)d->full; g.skysums = (const AS1 float *)d->skysums;
    g.hmax = (const AS4 float *)d->hmax; g.hmax_cols = (int)d->hmax_cols;
    g.gvf_len = (const AS1 uint32_t *)d->gvf_len; g.gvf_nosh = (const AS1 float *)d->gvf_nosh;
    g.march_trees = (const AS1 float *)d->march_trees;
    g.ptab_s = (const AS1 uint32_t *)d->ptab_s;
    g.gvf_edge = (const AS1 uint32_t *)d->gvf_edge;
    g.diffuse_tab = (const AS1 int8_t *)d->diffuse_tab;
    g.diffuse = (AS1 v2f *)d->diffuse;
    return g;
}

constexpr int CHUNK = 8;   // march / walk steps whose loads are issued together

// A non-negative integer-valued FP64 field of a uniform record (a table offset or length) as an
// int, with integer operations only, so the value stays in scalar registers (a float-to-int
// conversion would run on the vector unit and force the table addresses into vector registers).
__device__ __forceinline__ int uniform_index(double d) {
    const uint64_t b = (uint64_t)__double_as_longlong(d);
    const int e = (int)((b >> 52) & 0x7ffu) - 1023;
    if (e < 0) return 0;
    const uint64_t m = (b & ((1ull << 52) - 1)) | (1ull << 52);
    return (int)(m >> (52 - e));
}

__device__ __forceinline__ PixelAt locate_lane(const sw_domain *doms, const int32_t *bs_, int nd,
                                               int total, DomainG &G, int lane) {
    const AS4 int32_t *bs = (const AS4 int32_t *)bs_;
    PixelAt at;
    const int b = xcd_remap((int)blockIdx.x, total);
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bs[mid] <= b) lo = mid; else hi = mid - 1;
    }
    at.d = lo;
    G = domain_at(doms + at.d);
    const int t = b - bs[at.d];
    const int cols = (int)G.cols, rows = (int)G.rows;
    const int sh = G.tile_shift, tx = 1 << sh, ty = BLOCK >> sh;
    const int tpr = (cols + tx - 1) >> sh;
    const int tr0 = (t / tpr) * ty, tc0 = (t % tpr) * tx;
    at.r = tr0 + (lane >> sh);
    at.c = tc0 + (lane & (tx - 1));
    const int w64 = __builtin_amdgcn_readfirstlane(lane >> 6) * 64;
    at.wr0 = tr0 + (w64 >> sh);
    at.wr1 = tr0 + ((w64 + 63) >> sh);
    at.wc0 = tx > 64 ? tc0 + (w64 & (tx - 1)) : tc0;
    at.wc1 = tx > 64 ? at.wc0 + 63 : tc0 + tx - 1;
    at.tr0 = tr0; at.tc0 = tc0; at.tx = tx; at.ty = ty;
    at.tid = lane;
    at.valid = at.r < rows && at.c < cols;
    at.p = (long long)at.r * cols + at.c;
    return at;
}

__device__ __forceinline__ PixelAt locate(const sw_domain *doms, const int32_t *bs, int nd, int total, DomainG &G) {
    return locate_lane(doms, bs, nd, total, G, (int)threadIdx.x);
}

// FP64 heights of a pixel as series.py:252-287 forms them: a = dsm (+ raise), vegetation
// canopy DSM = trees + a and trunk-zone DSM = 0.25 trees + a, each zeroed where it equals a.
struct Heights {
    double a, top, bot;
};

template <typename PF>
__device__ __forceinline__ Heights heights_at(PF dsm, PF trees, double raise, long long q) {
#pragma clang fp contract(off)
    Heights h;
    h.a = (double)dsm[q] + raise;
    const double t = (double)trees[q];
    const double v1 = t + h.a;
    h.top = (v1 == h.a) ? 0.0 : v1;
    const double v2 = t * 0.25 + h.a;
    h.bot = (v2 == h.a) ? 0.0 : v2;
    return h;
}

__device__ __forceinline__ Heights heights_of(float dsm, float tree, double raise) {
#pragma clang fp contract(off)
    Heights h;
    h.a = (double)dsm + raise;
    const double t = (double)tree;
    const double v1 = t + h.a;
    h.top = (v1 == h.a) ? 0.0 : v1;
    const double v2 = t * 0.25 + h.a;
    h.bot = (v2 == h.a) ? 0.0 : v2;
    return h;
}

// =====================================================================================
// Static sky-view visibility: thread per (pixel, 32-patch word); geometry.cu:77-119.
// =====================================================================================
__global__ __launch_bounds__(BLOCK) void k_svf_visibility(
    const float *__restrict__ dsm, const float *__restrict__ trees, double raise, int rows,
    int cols, const int32_t *__restrict__ offsets, const double *__restrict__ heights,
    const int32_t *__restrict__ lengths, int capacity, int patches, uint32_t *__restrict__ bits,
    int tiles) {
#pragma clang fp contract(off)
    const int tpr = (cols + TILE_X - 1) / TILE_X;
    const int b = xcd_remap((int)blockIdx.x, tiles);
    const int r = (b / tpr) * TILE_Y + (int)threadIdx.x / TILE_X;
    const int c = (b % tpr) * TILE_X + (int)threadIdx.x % TILE_X;
    const int w = (int)blockIdx.y;
    if (r >= rows || c >= cols) return;
    const long long P = (long long)rows * cols;
    const long long p = (long long)r * cols + c;
    const Heights me = heights_at(dsm, trees, raise, p);
    const bool bushplant = (me.bot == 0.0) && me.top > 1.0;
    uint32_t wsh = 0, wveg = 0, wvb = 0;
    const int k0 = w * 32;
    const int k1 = min(patches, k0 + 32);
    for (int k = k0; k < k1; ++k) {
        const int32_t *off = offsets + (long long)k * capacity * 2;
        const double *dz = heights + (long long)k * capacity;
        const int L = lengths[k];
        bool sh = false, veg = bushplant, ever = false;
        for (int n = 0; n < L; ++n) {
            const int rr = r + off[2 * n], cc = c + off[2 * n + 1];
            if ((unsigned)rr < (unsigned)rows && (unsigned)cc < (unsigned)cols) {
                const Heights q = heights_at(dsm, trees, raise, (long long)rr * cols + cc);
                const double z = dz[n], zp = dz[n > 0 ? n - 1 : 0];
                if (q.a - z > me.a) { sh = true; veg = false; break; }
                const int cnt = (int)(q.top - z > me.a) + (int)(q.bot - z > me.a) +
                                (int)(q.top - zp > me.a) + (int)(q.bot - zp > me.a);
                veg = veg || (cnt > 0 && cnt < 4);
                ever = ever || veg;
            } else {
                sh = me.a < 0.0;
                if (sh) veg = false;
                ever = ever || veg;
                break;
            }
        }
        const uint32_t bit = 1u << (k - k0);
        if (!sh) wsh |= bit;                    // shmat = 1 - sh
        if (!veg) wveg |= bit;                  // vegmat = 1 - veg
        if (!(ever && !veg)) wvb |= bit;        // vbmat = 1 - (ever - veg)
    }
    bits[(0 * SW_VIS_WORDS + w) * P + p] = wsh;
    bits[(1 * SW_VIS_WORDS + w) * P + p] = wveg;
    bits[(2 * SW_VIS_WORDS + w) * P + p] = wvb;
}

// Band-weighted counts -> 15 SVF planes + svftotal (geometry.cu:121-159).
__global__ __launch_bounds__(256) void k_svf_reduce(
    const uint32_t *__restrict__ bits, const float *__restrict__ dsm,
    const float *__restrict__ trees, double raise, long long P, const uint8_t *__restrict__ sectors,
    const int32_t *__restrict__ band_offsets, const double *__restrict__ weights, int bands,
    double trans, double *__restrict__ out) {
#pragma clang fp contract(off)
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    uint32_t word[3][SW_VIS_WORDS];
    for (int m = 0; m < 3; ++m)
        for (int w = 0; w < SW_VIS_WORDS; ++w) word[m][w] = bits[(m * SW_VIS_WORDS + w) * P + p];
    double v[15];
    for (int i = 0; i < 15; ++i) v[i] = 0.0;
    for (int b = 0; b < bands; ++b) {
        int cnt[15];
        for (int i = 0; i < 15; ++i) cnt[i] = 0;
        for (int j = band_offsets[b]; j < band_offsets[b + 1]; ++j) {
            const uint8_t sec = sectors[j];
            for (int m = 0; m < 3; ++m) {
                const int on = (word[m][j >> 5] >> (j & 31)) & 1;
                cnt[m * 5] += on;
                for (int d = 0; d < 4; ++d)
                    if (sec & (1 << d)) cnt[m * 5 + 1 + d] += on;
            }
        }
        const double wf = weights[2 * b], wa = weights[2 * b + 1];
        for (int g = 0; g < 3; ++g) {
            v[g * 5] += wf * (double)cnt[g * 5];
            for (int d = 1; d < 5; ++d) v[g * 5 + d] += wa * (double)cnt[g * 5 + d];
        }
    }
    const double eps = 3.0459e-4;
    v[2] += eps;
    v[3] += eps;
    const Heights me = heights_at(dsm, trees, raise, p);
    const double last = me.bot == 0.0 ? eps : 0.0;
    v[7] += last; v[8] += last; v[12] += last; v[13] += last;
    for (int i = 0; i < 15; ++i) {
        if (v[i] > 1.0) v[i] = 1.0;
        out[i * P + p] = v[i];
    }
    out[15 * P + p] = v[0] - (1.0 - v[5]) * (1.0 - trans);
}

// The sine of the ground-temperature wave, sin(((frac - SNUP/24) / (TmaxLST/24 - SNUP/24)) pi/2)
// (umep_solweig.py:144-170): the hour record carries it in FP64 for the TmaxLST of SWH_TG_TMAX
// (the default maps' value); a pixel with another TmaxLST evaluates it in FP32.
__device__ __forceinline__ double tg_sine(const AS4 double *H, float tmaxlst) {
    if (tmaxlst == (float)H[SWH_TG_TMAX]) return H[SWH_TG_SINE];
    const float snup = (float)H[SWH_SNUP] * (1.0f / 24.0f);
    const float arg = ((float)H[SWH_FRAC] - snup) / (tmaxlst * (1.0f / 24.0f) - snup);
    return (double)sinf(arg * 1.57079632679489662f);
}

// =====================================================================================
// Per-hour sun-ray shadow march (shadowingfunction_wallheight_23, umep_solweig.py:2201-2406;
// geometry.cu:185-253) and the 16-byte record the ground-view walk gathers:
//   rec.x = Lup field of sunonsurface (umep_solweig.py:1418), rec.y = albedo, rec.z = albedo x
//   shadow, rec.w = flag word: byte 0 ground (0/1), byte 1 sunlit wall (0/1), byte 2 shadow
//   class (0: building shadow, 1: canopy shadow, 2: sunlit).
// =====================================================================================
__device__ __forceinline__ float shadow_of_class(uint32_t cls, float shv_canopy) {
    return cls == 2u ? 1.0f : (cls == 1u ? shv_canopy : 0.0f);
}

// Shadow class from the shadow pass's mask byte (bit 0: not building-shadowed, bit 1: no
// canopy shadow): 0 building shadow, 1 canopy shadow, 2 sunlit.
__device__ __forceinline__ uint32_t class_of(uint8_t m) {
    return (m & 1u) ? ((m & 2u) ? 2u : 1u) : 0u;
}

struct March {
    double f, topmax, prev;
    bool veg, ever;
};

constexpr int SKIP_WINDOW = 32;   // steps one test of the coarse height bound may skip
constexpr int SKIP_WINDOW_MAX = 64;   // after a skip the next test tries a window twice as long
constexpr int SKIP_TEST = 16;     // steps between tests (multiple of CHUNK)
static_assert(SKIP_WINDOW % SKIP_TEST == 0 && SKIP_WINDOW_MAX % SKIP_TEST == 0 && SKIP_TEST % CHUNK == 0,
              "tests stay on multiples of SKIP_TEST");
constexpr int HCELL_SHIFT = 4;    // 16 x 16 cells of the coarse height bound

extern "C" __device__ uint32_t __ockl_wfred_min_u32(uint32_t);

// A lower bound of the minimum over the wave's lanes of v >= 0 (every lane calls), as a uniform
// value: each v rounded down to FP32, whose bit patterns order as the values do, reduced with the
// DPP wave minimum. +inf stays +inf.
__device__ __forceinline__ double wave_min_below(double v) {
    const uint32_t b = __ockl_wfred_min_u32(__float_as_uint(__double2float_rd(v)));
    return (double)__uint_as_float(__builtin_amdgcn_readfirstlane(b));
}

// True when every sample of steps [k0, ke] of every lane of the wave is at most a_min above
// the ray: max over the 16 x 16 cells covering the samples' bounding box (the shifts grow
// monotonically, so the box spans the wave's pixels shifted by the first and the last step)
// minus the height of the step before k0. With a_min the smallest threshold of the wave's lanes,
// such steps change nothing: a - dz, top - dz, top - dz_prev and bot - dz all stay <= it.
__device__ __forceinline__ bool window_below(const DomainG &G, const PixelAt &at, const AS4 int32_t *off,
                                             int k0, int ke, int rows, int cols, double dz_before,
                                             double a_min) {
    const int dr0 = off[2 * k0], dr1 = off[2 * ke], dc0 = off[2 * k0 + 1], dc1 = off[2 * ke + 1];
    int r0 = at.wr0 + min(dr0, dr1), r1 = at.wr1 + max(dr0, dr1);
    int c0 = at.wc0 + min(dc0, dc1), c1 = at.wc1 + max(dc0, dc1);
    r0 = max(r0, 0); c0 = max(c0, 0); r1 = min(r1, rows - 1); c1 = min(c1, cols - 1);
    if (r0 > r1 || c0 > c1) return true;             // every sample is off the raster
    const int r0c = r0 >> HCELL_SHIFT, r1c = r1 >> HCELL_SHIFT, c0c = c0 >> HCELL_SHIFT, c1c = c1 >> HCELL_SHIFT;
    const int h = r1c - r0c + 1, w = c1c - c0c + 1, hc = G.hmax_cols;
    const long long level = (long long)((rows + (1 << HCELL_SHIFT) - 1) >> HCELL_SHIFT) * hc;
    float hm;
    if (h <= 8 && w <= 8) {
        // four blocks of 2^i x 2^j cells (i, j the floor log2 of the box's sides, at most 2)
        // cover the box exactly: four independent scalar loads
        const int i = h >= 4 ? 2 : (h >= 2 ? 1 : 0), j = w >= 4 ? 2 : (w >= 2 ? 1 : 0);
        const AS4 float *M = G.hmax + (i * 3 + j) * level;
        const int ra = r0c * hc, rb = (r1c - (1 << i) + 1) * hc, cb = c1c - (1 << j) + 1;
        hm = fmaxf(fmaxf(M[ra + c0c], M[ra + cb]), fmaxf(M[rb + c0c], M[rb + cb]));
    } else {
        hm = -__builtin_huge_valf();
        for (int cr = r0c; cr <= r1c; ++cr)
            for (int cc = c0c; cc <= c1c; ++cc) hm = fmaxf(hm, G.hmax[cr * hc + cc]);
    }
    return (double)hm - dz_before <= a_min;
}

typedef int32_t i32x16 __attribute__((ext_vector_type(16)));
typedef double f64x8 __attribute__((ext_vector_type(8)));

// CHUNK steps [k0, k0 + CHUNK) of the sun-ray march of one pixel. MASKED: per-lane bounds
// tests (a step outside the raster is made a no-op by an infinite dz) and only steps < L count;
// otherwise all CHUNK steps are inside the raster for every lane and below L. TOPS: ``trees`` is
// the march plane, -inf where a sample has no canopy top: top = t + a and bot = 0.25 t + a are
// then -inf there, which leaves topmax (>= 0), the pergola test and the receiver threshold
// (>= 0, against dz >= 0) exactly as the zeroed tops of the reference do.
// A buffer resource over one FP32 plane of n values (the march's gathers: a 32-bit byte offset per
// lane instead of a 64-bit address, shared by the DSM and canopy planes of one layout; loads past
// the plane return 0 instead of faulting).
typedef __amdgpu_buffer_rsrc_t PlaneRsrc;
__device__ __forceinline__ PlaneRsrc plane_rsrc(const AS1 float *plane, long long n) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)(uintptr_t)plane, (short)0, (int)(n * 4), 0x00020000);
}

// SKIP_BLOCKED: the caller needs no canopy terms (topmax, the pergola test) of a lane that is
// building-shadowed when the chunk starts, nor of a lane past the raster: veg is false for good
// there, ever keeps its value, and topmax enters the outputs only through veg.
template <bool MASKED, bool TOPS, bool SKIP_BLOCKED>
__device__ __forceinline__ void march_chunk(March &m, PlaneRsrc dsm, PlaneRsrc trees,
                                            double raise, double me_a, const PixelAt &at, int rows,
                                            int cols, const AS4 int32_t *off, const AS4 double *dzs,
                                            int k0, int L) {
#pragma clang fp contract(off)
    static_assert(CHUNK == 8, "one 16-dword load of shifts and one of heights per chunk");
    float av[CHUNK], tv[CHUNK];
    bool in[CHUNK];
    // the chunk's 8 shift pairs and 8 heights as two scalar loads (ray slices start at multiples
    // of 8 steps)
    const i32x16 ov = *(const AS4 i32x16 *)(off + 2 * k0);
    const f64x8 zv = *(const AS4 f64x8 *)(dzs + k0);
    const uint32_t p4 = (uint32_t)at.p * 4u;
#pragma unroll
    for (int i = 0; i < CHUNK; ++i) {
        const int dr = ov[2 * i], dc = ov[2 * i + 1];
        uint32_t o;
        if (!MASKED) {
            in[i] = true;
            o = p4 + (uint32_t)(dr * cols + dc) * 4u;
        } else {
            // a lane outside the raster reads its own pixel: selects, not branches, so that the
            // chunk stays one block and its 8 offset pairs load together
            const int rr = at.r + dr, cc = at.c + dc;
            in[i] = (unsigned)rr < (unsigned)rows && (unsigned)cc < (unsigned)cols;
            const int rq = in[i] ? rr : at.r, cq = in[i] ? cc : at.c;
            o = (uint32_t)(rq * cols + cq) * 4u;
        }
        av[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(dsm, (int)o, 0, 0));
        tv[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(trees, (int)o, 0, 0));
    }
    if (TOPS) {
        // no canopy top among the chunk's samples of any lane: top and bot are -inf, so topmax
        // keeps its value and the pergola test fails at every step; only f, veg and ever move.
        // The march plane holds -inf or a top height > 0: the chunk has a top exactly when the
        // largest sample is not -inf (a chain of three-operand maxima instead of a flag per
        // sample; the same test on the integer bit patterns holds 7 more registers).
        float tmax = tv[0];
#pragma unroll
        for (int i = 1; i < CHUNK; ++i) tmax = __builtin_fmaxf(tmax, tv[i]);
        const bool open = !SKIP_BLOCKED || (at.valid && !(m.f > me_a));
        if (__ballot(open && tmax != -__builtin_huge_valf()) == 0) {
#pragma unroll
            for (int i = 0; i < CHUNK; ++i) {
                const bool live = !MASKED || (in[i] && k0 + i < L);
                const double z = zv[i];
                const double dz = live ? z : __builtin_huge_val();
                const double a = (double)av[i] + raise;
                m.f = fmax(m.f, a - dz);
                m.prev = z;
                m.veg = m.veg && !(m.f > me_a);
                m.ever = m.ever || m.veg;
            }
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < CHUNK; ++i) {
        // a step outside the raster or past the ray's end (the padding) changes nothing: its dz
        // is +inf, so the chunk has no branches and its loads issue together
        const bool live = !MASKED || (in[i] && k0 + i < L);
        const double z = zv[i];
        const double dz = live ? z : __builtin_huge_val();
        const double dzp = live ? m.prev : __builtin_huge_val();
        const double a = (double)av[i] + raise;
        const double t = (double)tv[i];
        const double v1 = t + a;
        const double v2 = __builtin_fma(t, 0.25, a);   // t * 0.25 is exact: one rounding, as NumPy
        const double top = TOPS ? v1 : ((v1 == a) ? 0.0 : v1);
        const double bot = TOPS ? v2 : ((v2 == a) ? 0.0 : v2);
        m.f = fmax(m.f, a - dz);
        m.topmax = fmax(m.topmax, top - dz);
        const bool vstep = (top - dzp > me_a) && !(bot - dz > me_a);
        m.prev = z;
        m.veg = (m.veg || vstep) && !(m.f > me_a);
        m.ever = m.ever || m.veg;
    }
}

// Hour hour0 + blockIdx.y. With per_hour the hour's record and mask byte go to slot ``hour`` of
// the domain's rec / masks planes ([hours][P]), else to their single plane.
template <bool FULL>
__global__ __launch_bounds__(BLOCK) void k_hour_shadow(const sw_domain *__restrict__ doms,
                                                       const int32_t *__restrict__ bs, int nd,
                                                       int total, int hour0, int per_hour) {
#pragma clang fp contract(off)
    DomainG G;
    PixelAt at = locate(doms, bs, nd, total, G);
    const int hour = hour0 + (int)blockIdx.y;
    const AS4 double *H = G.hours + (long long)hour * SW_HOUR_STRIDE;
    if (H[SWH_DAY] == 0.0) return;
    if (per_hour) {
        G.rec += (long long)hour * G.pixels;
        if (G.masks) G.masks += (long long)hour * G.pixels;
    }
    const int rows = (int)G.rows, cols = (int)G.cols;
    // Lanes past the raster (on the last tile row or column) march along with a copy of the
    // nearest pixel, so that the wave-wide thresholds below see every lane; they neither limit a
    // skip nor keep the wave marching, and write nothing.
    const bool valid = at.valid;
    if (!valid) {
        at.r = min(at.r, rows - 1);
        at.c = min(at.c, cols - 1);
        at.p = (long long)at.r * cols + at.c;
    }
    const long long P = G.pixels, p = at.p;
    const AS1 float *dsm = G.dsm, *trees = G.trees;
    const double raise = G.dsm_raise;
    const Heights me = heights_at(dsm, trees, raise, p);
    const bool tops = G.march_trees != nullptr;       // uniform: one domain per workgroup
    const AS1 float *mtrees_ = tops ? G.march_trees : trees;
    const PlaneRsrc rdsm = plane_rsrc(dsm, P), mtrees = plane_rsrc(mtrees_, P);
    const bool wall_px = valid && G.walls[p] > 0.0f;
    const double wall_h = (double)G.walls[p];
    const int start = uniform_index(H[SWH_RAY_START]);
    const int L = uniform_index(H[SWH_RAY_LEN]);
    const AS4 int32_t *off = G.ray_off + 2LL * start;
    const AS4 double *dzs = G.ray_dz + start;
    // The march follows the reference's array form: f = running max of a - dz, the canopy
    // pergola test on (top, bot) x (dz, previous dz), and topmax of top - dz. Because trees >= 0,
    // bot <= top, so of the four canopy differences B = bot - dz is the smallest and Tp = top - prev
    // the largest: "1 to 3 of the 4 above the receiver" is Tp > a && !(B > a).
    // A shift that leaves the raster leaves it for good (row and column shifts grow in magnitude
    // with the step), and such a step changes nothing (the zero fill of the shifted arrays is never
    // above a >= 0), so each chunk of steps is either inside the raster for every lane of the wave
    // (no bounds tests) or is evaluated with per-lane tests; the wave stops when no lane is inside.
    March m;
    m.f = me.a; m.topmax = me.top; m.prev = 0.0;
    m.veg = (me.bot == 0.0) && me.top > 1.0;
    m.ever = false;
    const bool all_lanes = true;
    const uint64_t active = __ballot(all_lanes);
    int win = SKIP_WINDOW;                            // uniform: doubled after a skip, reset after a miss
    for (int k0 = 0; k0 < L; k0 += CHUNK) {
        if (k0 > 0 && (k0 & (SKIP_TEST - 1)) == 0) {
            // A lane's remaining steps matter only through samples above its threshold: while
            // not building-shadowed, its own height (f, the canopy test and topmax change only
            // through samples above it); once shadowed, veg is false for good (so ever and
            // topmax no longer matter), and the running maximum f matters only on a wall pixel:
            // for the FP64 wall planes (FULL), else only while wall - (f - a) still rounds to
            // the wall height (a sunlit wall needs wallsun == wall, and f only grows).
            // A window whose samples stay at or below every lane's threshold changes nothing;
            // when no lane has a threshold left, the march is over.
            const bool wall_open = wall_px && (FULL || wall_h - (m.f - me.a) == wall_h);
            const double thr = !valid ? __builtin_huge_val()
                                      : (m.f > me.a ? (wall_open ? m.f : __builtin_huge_val()) : me.a);
            const double thr_min = wave_min_below(thr);
            if (thr_min == __builtin_huge_val()) break;
            const int ke = (k0 + win < L ? k0 + win : L) - 1;
            if (window_below(G, at, off, k0, ke, rows, cols, dzs[k0 - 1], thr_min)) {
                m.prev = dzs[ke];
                k0 = ke + 1 - CHUNK;
                win = win < SKIP_WINDOW_MAX ? 2 * win : SKIP_WINDOW_MAX;
                continue;
            }
            win = SKIP_WINDOW;
        }
        const int kl = k0 + CHUNK - 1 < L ? k0 + CHUNK - 1 : L - 1;
        const bool in_last = (unsigned)(at.r + off[2 * kl]) < (unsigned)rows &&
                             (unsigned)(at.c + off[2 * kl + 1]) < (unsigned)cols;
        const bool in_first = valid && (unsigned)(at.r + off[2 * k0]) < (unsigned)rows &&
                              (unsigned)(at.c + off[2 * k0 + 1]) < (unsigned)cols;
        if (__ballot(in_first) == 0) break;          // every lane's ray has left the raster
        const bool full = k0 + CHUNK <= L && __ballot(in_last) == active;
        if (full && tops)
            march_chunk<false, true, !FULL>(m, rdsm, mtrees, raise, me.a, at, rows, cols, off, dzs, k0, L);
        else if (full)
            march_chunk<false, false, !FULL>(m, rdsm, mtrees, raise, me.a, at, rows, cols, off, dzs, k0, L);
        else if (tops)
            march_chunk<true, true, !FULL>(m, rdsm, mtrees, raise, me.a, at, rows, cols, off, dzs, k0, L);
        else
            march_chunk<true, false, !FULL>(m, rdsm, mtrees, raise, me.a, at, rows, cols, off, dzs, k0, L);
    }
    if (!valid) return;
    const double f = m.f, topmax = m.topmax;
    const bool veg = m.veg, ever = m.ever;
    const bool blocked = f > me.a;
    const double wall = (double)G.walls[p];
    const double wallbol = wall > 0.0 ? 1.0 : 0.0;
    const double aspect = ((double)G.aspect[p] * 3.141592653589793) / 180.0;
    const double low = H[SWH_FACE_LOW], high = H[SWH_FACE_HIGH];
    double facesh;
    if (H[SWH_FACE_MIDDLE] != 0.0)
        facesh = (double)(aspect < low || aspect >= high) - wallbol + 1.0;
    else
        facesh = 1.0 - (double)(aspect > low || aspect <= high);
    const double vegv = veg ? 1.0 : 0.0;
    const double shvoveg = (topmax - me.a) * vegv;
    double wallsun = wall - (f - me.a);
    if (wallsun < 0.0) wallsun = 0.0;
    if (facesh == 1.0) wallsun = 0.0;
    const double wallsh = wall - wallsun;
    double wallshve = shvoveg * wallbol - wallsh;
    if (wallshve < 0.0) wallshve = 0.0;
    if (wallshve > wall) wallshve = wall;
    wallsun = wallsun - wallshve;
    if (wallsun < 0.0) { wallshve = 0.0; wallsun = 0.0; }
    const double sh = blocked ? 0.0 : 1.0;
    const double vegsh = 1.0 - vegv;
    const double shadow = sh - (1.0 - vegsh) * (1.0 - H[SWH_PSI]);
    const bool ground = G.ground[p] != 0;
    // (wallsun / walls * buildings) == 1 (umep_solweig.py:606): with wall > 0 a correctly rounded
    // quotient is 1 exactly when wallsun == wall; with wall <= 0, wallsun is 0 or -0 and the
    // quotient is -0 or NaN
    const bool sunwall = ground && wall > 0.0 && wallsun == wall;
    // Surface temperature of the hour (umep_solweig.py:144-170) and the sunonsurface Lup field.
    const double amp = (double)G.tgk[p] * H[SWH_ALTMAX] + (double)G.tstart[p];
    double tg = amp * tg_sine(H, G.tmaxlst[p]);
    tg = tg * H[SWH_CITGG];
    if (G.landcover == 1 && tg < 0.0) tg = 0.0;
    const double ta = H[SWH_TA], emis = (double)G.emis[p];
    const double tk = tg * shadow + ta + 273.15, ta4 = (ta + 273.15);
    const double lup = SBC * emis * (tk * tk) * (tk * tk) - SBC * emis * (ta4 * ta4) * (ta4 * ta4);
    const uint32_t cls = blocked ? 0u : (veg ? 1u : 2u);
    const float albsh = G.alb[p] * shadow_of_class(cls, (float)(1.0 - (1.0 - H[SWH_PSI])));
    v2f rec;
    rec.x = (float)lup;
    rec.y = __uint_as_float(__float_as_uint(albsh) | (sunwall ? 0x80000000u : 0u));
    G.rec[p] = rec;
    if (G.masks)
        G.masks[p] = (uint8_t)((blocked ? 0 : 1) | (veg ? 0 : 2) | ((ever && !veg) ? 0 : 4) |
                                (sunwall ? 8 : 0) | (facesh == 1.0 ? 16 : 0));
    if (FULL && G.shadow_planes) {
        AS1 double *o = G.shadow_planes;
        const double facesun = ((facesh + wallbol == 1.0) && wall > 0.0) ? 1.0 : 0.0;
        o[p] = vegsh;
        o[P + p] = sh;
        o[2 * P + p] = 1.0 - ((ever ? 1.0 : 0.0) - vegv);
        o[3 * P + p] = wallsh;
        o[4 * P + p] = wallsun;
        o[5 * P + p] = wallshve;
        o[6 * P + p] = facesh;
        o[7 * P + p] = facesun;
    }
}

// Static per-pixel patch sums of the hour-independent longwave terms (from the visibility bits).
// Static sky-view visibility as the sun-ray march of the hour kernel (geometry.cu:77-119): a patch is
// seen from a pixel exactly when a sun at the patch centre would light it, so every patch's ray
// (svf.py:19-78, step 0 included) is marched with march_chunk: 8-step chunks without bounds tests
// while the whole wave stays inside the raster, a window of 32 (then 64) steps skipped when no
// 16 x 16 height-bound cell it crosses rises above the wave's lowest live threshold (a pixel's own
// height until a building blocks it: the reference stops there, its canopy flag cleared), and the
// patch ends when every lane is blocked or has left the raster. An off-raster step changes nothing
// for heights >= 0 (the reference's sh = a < 0 exit), which the caller guarantees (dsm + raise).
// One wave holds a 16 x 4 block of a 16 x 16 tile and marches its pixels through the 32 patches of
// blockIdx.y's visibility word.
__global__ __launch_bounds__(BLOCK) void k_svf_march(const float *__restrict__ dsm_, const float *__restrict__ trees_,
                                                     const float *__restrict__ mtrees_, const float *__restrict__ hmax_,
                                                     int hmax_cols, double raise, int rows, int cols,
                                                     const int32_t *__restrict__ offsets_,
                                                     const double *__restrict__ heights_,
                                                     const int32_t *__restrict__ lengths_, int capacity, int patches,
                                                     uint32_t *__restrict__ bits, int tiles) {
#pragma clang fp contract(off)
    constexpr int TX = 16, SH = 4;
    DomainG G = {};
    G.rows = rows; G.cols = cols; G.pixels = (long long)rows * cols; G.dsm_raise = raise; G.tile_shift = SH;
    G.hmax = (const AS4 float *)hmax_; G.hmax_cols = hmax_cols;
    PixelAt at;
    const int tpr = (cols + TX - 1) >> SH;
    const int b = xcd_remap((int)blockIdx.x, tiles);
    at.tr0 = (b / tpr) * (BLOCK / TX);
    at.tc0 = (b % tpr) * TX;
    at.tx = TX; at.ty = BLOCK / TX;
    at.r = at.tr0 + ((int)threadIdx.x >> SH);
    at.c = at.tc0 + ((int)threadIdx.x & (TX - 1));
    const int w64 = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6) * 64;
    at.wr0 = at.tr0 + (w64 >> SH);
    at.wr1 = at.tr0 + ((w64 + 63) >> SH);
    at.wc0 = at.tc0;
    at.wc1 = at.tc0 + TX - 1;
    // lanes past the raster march with a copy of the nearest pixel (so the wave-wide thresholds see
    // every lane) and write nothing
    const bool valid = at.r < rows && at.c < cols;
    at.valid = valid;
    if (!valid) {
        at.r = min(at.r, rows - 1);
        at.c = min(at.c, cols - 1);
    }
    at.p = (long long)at.r * cols + at.c;
    const long long P = G.pixels, p = at.p;
    const AS1 float *dsm = (const AS1 float *)dsm_, *trees = (const AS1 float *)trees_;
    const bool tops = mtrees_ != nullptr;
    const PlaneRsrc rdsm = plane_rsrc(dsm, G.pixels),
                    mtrees = plane_rsrc(tops ? (const AS1 float *)mtrees_ : trees, G.pixels);
    const Heights me = heights_at(dsm, trees, raise, p);
    const bool bushplant = (me.bot == 0.0) && me.top > 1.0;
    const AS4 int32_t *lens = (const AS4 int32_t *)lengths_;
    const int w = (int)blockIdx.y;
    const int kb = w * 32, ke_patch = min(patches, kb + 32);
    const uint64_t active = __ballot(true);
    uint32_t wsh = 0u, wveg = 0u, wvb = 0u;
    for (int k = kb; k < ke_patch; ++k) {
        const AS4 int32_t *off = (const AS4 int32_t *)offsets_ + (long long)k * capacity * 2;
        const AS4 double *dzs = (const AS4 double *)heights_ + (long long)k * capacity;
        const int L = lens[k];
        March m;
        m.f = me.a; m.topmax = me.top; m.prev = 0.0;
        m.veg = bushplant;
        m.ever = false;
        int win = SKIP_WINDOW;
        for (int k0 = 0; k0 < L; k0 += CHUNK) {
            if (k0 > 0 && (k0 & (SKIP_TEST - 1)) == 0) {
                const double thr = (!valid || m.f > me.a) ? __builtin_huge_val() : me.a;
                const double thr_min = wave_min_below(thr);
                if (thr_min == __builtin_huge_val()) break;       // every lane blocked
                const int kl = (k0 + win < L ? k0 + win : L) - 1;
                if (window_below(G, at, off, k0, kl, rows, cols, dzs[k0 - 1], thr_min)) {
                    m.prev = dzs[kl];
                    k0 = kl + 1 - CHUNK;
                    win = win < SKIP_WINDOW_MAX ? 2 * win : SKIP_WINDOW_MAX;
                    continue;
                }
                win = SKIP_WINDOW;
            }
            const int kl = k0 + CHUNK - 1 < L ? k0 + CHUNK - 1 : L - 1;
            const bool in_last = (unsigned)(at.r + off[2 * kl]) < (unsigned)rows &&
                                 (unsigned)(at.c + off[2 * kl + 1]) < (unsigned)cols;
            const bool in_first = valid && (unsigned)(at.r + off[2 * k0]) < (unsigned)rows &&
                                  (unsigned)(at.c + off[2 * k0 + 1]) < (unsigned)cols;
            if (__ballot(in_first) == 0) break;              // every lane's ray has left the raster
            const bool full = k0 + CHUNK <= L && __ballot(in_last) == active;
            if (full && tops)
                march_chunk<false, true, true>(m, rdsm, mtrees, raise, me.a, at, rows, cols, off, dzs, k0, L);
            else if (full)
                march_chunk<false, false, true>(m, rdsm, mtrees, raise, me.a, at, rows, cols, off, dzs, k0, L);
            else if (tops)
                march_chunk<true, true, true>(m, rdsm, mtrees, raise, me.a, at, rows, cols, off, dzs, k0, L);
            else
                march_chunk<true, false, true>(m, rdsm, mtrees, raise, me.a, at, rows, cols, off, dzs, k0, L);
        }
        const uint32_t bit = 1u << (k - kb);
        const bool sh = m.f > me.a;
        if (!sh) wsh |= bit;                                 // shmat = 1 - sh
        if (!m.veg) wveg |= bit;                             // vegmat = 1 - veg
        if (!(m.ever && !m.veg)) wvb |= bit;                 // vbmat = 1 - (ever - veg)
    }
    if (!valid) return;
    bits[(0 * SW_VIS_WORDS + w) * P + p] = wsh;
    bits[(1 * SW_VIS_WORDS + w) * P + p] = wveg;
    bits[(2 * SW_VIS_WORDS + w) * P + p] = wvb;
}

__global__ __launch_bounds__(256) void k_sky_sums(const uint32_t *__restrict__ bits, long long P,
                                                  const float *__restrict__ tab, float *__restrict__ out) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    float s[SW_SKY_SUMS];
    for (int i = 0; i < SW_SKY_SUMS; ++i) s[i] = 0.f;
    for (int w = 0; w < SW_VIS_WORDS; ++w) {
        const uint32_t vsh = bits[(0 * SW_VIS_WORDS + w) * P + p];
        const uint32_t vveg = bits[(1 * SW_VIS_WORDS + w) * P + p];
        const uint32_t vvb = bits[(2 * SW_VIS_WORDS + w) * P + p];
        const uint32_t sky = vsh & vveg, vegf = ~vveg | ~vvb, bld = ~vsh & vvb, obst = ~vsh | ~vveg | ~vvb;
        const int k0 = w * 32, k1 = min(SW_PATCHES, k0 + 32);
        for (int k = k0; k < k1; ++k) {
            const int j = k - k0;
            const float *c = tab + 4 * k;
            if ((sky >> j) & 1) { s[SWS_SKY_SIN] += c[0]; s[SWS_SKY_COS] += c[1]; s[SWS_SKY_GSIN] += c[2]; s[SWS_SKY_GCOS] += c[3]; }
            if ((vegf >> j) & 1) { s[SWS_VEG_SIN] += c[0]; s[SWS_VEG_COS] += c[1]; }
            if ((obst >> j) & 1) { s[SWS_OBS_SIN] += c[0]; s[SWS_OBS_COS] += c[1]; }
            if ((bld >> j) & 1) { s[SWS_BLD_SIN] += c[0]; s[SWS_BLD_COS] += c[1]; }
        }
    }
    for (int i = 0; i < SW_SKY_SUMS; ++i) out[i * P + p] = s[i];
}

// =====================================================================================
// Per-hour pixel pass: ground view factors (sunonsurface_2018a / gvf_2018a), lag recursion,
// wedge, Kdown/Kup, anisotropic Kside and longwave patch sums, Tmrt, UTCI.
// =====================================================================================
struct Triple {
    float lup, alb, nosh;
};

// influence (near or far field): with a sunlit / any wall in the field, walls and ground share
// the divisor ``n + 1``, else the ground alone is divided by ``n`` (gvf.cu:209-239); divisors
// enter as reciprocals.
__device__ __forceinline__ float influence(float wall_sum, float ground_sum, float inv_wall_div,
                                           float inv_ground_div, bool has_wall) {
    return has_wall ? (wall_sum + ground_sum) * inv_wall_div : ground_sum * inv_ground_div;
}

// (near * 0.5 + far * 0.4) / 0.9 as two products
constexpr float NEAR_W = 0.5f / 0.9f, FAR_W = 0.4f / 0.9f;
__device__ __forceinline__ float distance_weight(float near_v, float far_v) {
    return near_v * NEAR_W + far_v * FAR_W;
}

constexpr float INV18 = 1.0f / 18.0f, INV9 = 1.0f / 9.0f;

struct GvfOut {
    Triple all, e, s, w, n;
};

// Running state of one direction of the adjoint's walk. ``af`` is 1 while the running minimum of
// the ground mask is 1 (umep_solweig.py:1498) and 0 from the first non-ground sample on; every
// term is multiplied by it, so the walk needs no per-lane branches. ``L`` counts the ground
// steps, ``wf`` is 1 once a sunlit wall was passed, ``cw`` counts the ground steps from there.
struct Walk {
    float af, lup, alb, albsh, L, wf, cw;
};

// One pixel's walk over the 18 directions x ``second`` steps (sunonsurface_2018a). The number
// of steps before the first non-ground sample, L, depends only on the ground mask and is a static
// plane (gvf_len); the walk sums the hour's Lup and albedo x shadow over those L steps and keeps
// a bit per step where the sample is a sunlit wall (the record's sign bit). Past step L every
// step adds one "any wall" count and, once a sunlit wall was passed, one sunlit-wall count
// (gvf.cu:209-239 sums the same terms step by step): with k the first sunlit-wall step, the far
// field counts second - k sunlit-wall steps and second - L wall steps, the near field (the first
// ``first`` steps) first - k and first - min(L, first). The unshadowed-albedo sums do not depend
// on the hour and come from the static planes (gvf_nosh). On the domain edge a step whose shift
// leaves the raster repeats the previous sample, as the shifted-window assignment of the
// reference leaves it.
// The hour records of a workgroup's tile and its ground-view halo, staged in LDS (rows of
// ``pitch`` records) when the whole halo lies inside the raster.
extern __shared__ v2f gv_lds[];

// Copy the tile's records and a halo of h pixels on every side into LDS (all threads of the
// workgroup take part; the caller has checked that the region lies inside the raster).
__device__ __forceinline__ void stage_tile(const DomainG &G, const PixelAt &at, int h, int pitch) {
    const int cols = (int)G.cols, nrows = at.ty + 2 * h;
    const int wave = at.tid >> 6, lane = at.tid & 63;
    const AS1 v2f *rec = G.rec;
    for (int rr = wave; rr < nrows; rr += BLOCK / 64) {
        const long long src = (long long)(at.tr0 - h + rr) * cols + (at.tc0 - h);
        for (int cc = lane; cc < pitch; cc += 64) gv_lds[rr * pitch + cc] = rec[src + cc];
    }
    __syncthreads();
}

// The per-direction sums of one pixel's walk folded into the four halves E (azimuth < 180),
// S (90 <= azimuth < 270), W and N; the sum over all 18 directions is E + W. A sunlit wall in
// the far field (kw < L) or the near field (kw < min(L, first)) adds (n - kw) wall terms and
// switches the field's divisor from n to n + 1 (gvf.cu:209-239); ``keep`` removes the far-field
// wall terms of a receiver that is itself a sunlit wall facing the direction (facesh == 0, a
// static bit of gvf_len). Each direction's Lup and albedo terms are then linear in the wall flux
// (lwall) and wall albedo (albedo_b) with one weight B. With first == 1 the near field is the
// receiver alone (L >= 1 exactly on ground pixels, kw == 0 exactly on a sunlit-wall receiver),
// the same for every direction: it is added once per half by finish().
struct GvSums {
    float firstf, secondf, kF1, kF0, kN1, kN0;
    // wall weight, Lup and albedo x shadow sums of the quadrants NE (azimuth < 90), SE, SW and NW;
    // each half is the sum of its two quadrants (E = NE + SE, S = SE + SW, W = SW + NW, N = NW + NE)
    float hB[4], hL[4], hA[4];

    __device__ __forceinline__ GvSums(int first, int second) {
        firstf = (float)first;
        secondf = (float)second;
        // per direction weights of the near (first steps) and far (all steps) fields, with and
        // without a sunlit wall in the field (the divisors n + 1 and n)
        kF1 = FAR_W * (1.0f / (secondf + 1.0f));
        kF0 = FAR_W * (1.0f / secondf);
        kN1 = NEAR_W * (1.0f / (firstf + 1.0f));
        kN0 = NEAR_W * (1.0f / firstf);
#pragma unroll
        for (int g = 0; g < 4; ++g) hB[g] = hL[g] = hA[g] = 0.f;
    }

    __device__ __forceinline__ void add_to(int dir, float B, float tl, float ta) {
        const int q = (5 + 20 * dir) / 90;
        // a constant quadrant in the unrolled walks; the rolled walks select (no indexed registers)
#pragma unroll
        for (int g = 0; g < 4; ++g)
            if (g == q) { hB[g] += B; hL[g] += tl; hA[g] += ta; }
    }

    // far field of one direction: lup and albsh are the sums over its L alive steps, kw the
    // first sunlit-wall step of the walk (any value >= L when none lies among the alive steps)
    __device__ __forceinline__ void add_far(int dir, int L, uint32_t lb, uint32_t kw, float lup, float albsh) {
        const bool sunin2 = kw < (uint32_t)L;
        const bool keep = kw == 0u && (lb & 0x80u) != 0u;
        const float kF = sunin2 ? kF1 : kF0;
        const float wF = (sunin2 && !keep) ? secondf - (float)kw : 0.f;
        add_to(dir, kF * wF, kF * lup, kF * albsh);
    }

    // near field of one direction when first > 1: sums over its first min(L, first) steps
    __device__ __forceinline__ void add_near(int dir, int L, uint32_t kw, int first, float f_lup, float f_albsh) {
        const uint32_t Lf = (uint32_t)(L < first ? L : first);
        const bool f_sunin = kw < Lf;
        const float kN = f_sunin ? kN1 : kN0;
        const float wN = f_sunin ? firstf - (float)kw : 0.f;
        add_to(dir, kN * wN, kN * f_lup, kN * f_albsh);
    }

    // the receiver's own terms (n x its Lup and albedo x shadow on non-building pixels), the
    // uniform near field when first == 1, and the hour-independent unshadowed-albedo sums of the
    // static planes (gvf_nosh)
    __device__ __forceinline__ GvfOut finish(const DomainG &G, const PixelAt &at, const AS4 double *H,
                                             const v2f me, float building, bool near_uniform) {
        const float lwall = (float)H[SWH_LWALL];
        const float albedo_b = (float)H[SWH_ALBEDO_B];
        const float local_lup = me.x * (1.0f - building);
        const float local_albsh = fabsf(me.y) * (1.0f - building);
        // the halves E, S, W, N from the quadrants
        float B[4], Lh[4], Ah[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            B[g] = hB[g] + hB[(g + 1) & 3];
            Lh[g] = hL[g] + hL[(g + 1) & 3];
            Ah[g] = hA[g] + hA[(g + 1) & 3];
        }
        if (near_uniform) {
            // receiver on the ground (L >= 1) and a sunlit wall (its record's sign bit): kw == 0
            const bool f_sunin = building != 0.0f && (__float_as_uint(me.y) >> 31) != 0u;
            const float kN = f_sunin ? kN1 : kN0;
            const float nB = 9.0f * (f_sunin ? kN : 0.0f);
            const float nL = 9.0f * (kN * (building * me.x));
            const float nA = 9.0f * (kN * (building * fabsf(me.y)));
#pragma unroll
            for (int g = 0; g < 4; ++g) { B[g] += nB; Lh[g] += nL; Ah[g] += nA; }
        }
        GvfOut o;
        Triple *grp[4] = {&o.e, &o.s, &o.w, &o.n};
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            grp[g]->lup = fmaf(lwall, B[g], Lh[g]) + 9.0f * local_lup;
            grp[g]->alb = fmaf(albedo_b, B[g], Ah[g]) + 9.0f * local_albsh;
        }
        o.all.lup = fmaf(lwall, B[0] + B[2], Lh[0] + Lh[2]) + 18.0f * local_lup;
        o.all.alb = fmaf(albedo_b, B[0] + B[2], Ah[0] + Ah[2]) + 18.0f * local_albsh;
        const AS1 float *ns = G.gvf_nosh;
        const long long P = G.pixels, p = at.p;
        o.all.nosh = fmaf(albedo_b, ns[5 * P + p], ns[p]);
        o.e.nosh = fmaf(albedo_b, ns[6 * P + p], ns[P + p]);
        o.s.nosh = fmaf(albedo_b, ns[7 * P + p], ns[2 * P + p]);
        o.w.nosh = fmaf(albedo_b, ns[8 * P + p], ns[3 * P + p]);
        o.n.nosh = fmaf(albedo_b, ns[9 * P + p], ns[4 * P + p]);
        return o;
    }
};

// stage_tile for a TX-wide tile and a halo of HH pixels: every load of a thread is issued before
// the first LDS store (the loop form waits for each row's load in turn). Wave w copies rows
// w, w + 4, ... of the region, its lanes the row's records; lanes past the region's last row or
// column copy that row or column again (the same value to the same place), so that no load or
// store is conditional.
// CLAMP (a tile whose halo leaves the raster): positions outside the raster take the nearest
// raster record; the walk never uses them (ground_view_fixed<TX, true> repeats the last in-raster
// sample instead).
template <int TX, int HH, bool CLAMP>
__device__ __forceinline__ void stage_tile_fixed(const DomainG &G, const PixelAt &at) {
    constexpr int PITCH = TX + 2 * HH, NROWS = BLOCK / TX + 2 * HH;
    constexpr int WAVES = BLOCK / 64, ROWS = (NROWS + WAVES - 1) / WAVES, PASSES = (PITCH + 63) / 64;
    const int rows = (int)G.rows, cols = (int)G.cols, wave = at.tid >> 6, lane = at.tid & 63;
    const AS1 v2f *rec = G.rec + (CLAMP ? 0LL : ((long long)(at.tr0 - HH) * cols + (at.tc0 - HH)));
    v2f v[ROWS][PASSES];
#pragma unroll
    for (int k = 0; k < ROWS; ++k)
#pragma unroll
        for (int j = 0; j < PASSES; ++j) {
            const int rr = min(wave + WAVES * k, NROWS - 1), cc = min(lane + 64 * j, PITCH - 1);
            if (CLAMP) {
                const int sr = min(max(at.tr0 - HH + rr, 0), rows - 1), sc = min(max(at.tc0 - HH + cc, 0), cols - 1);
                v[k][j] = rec[(long long)sr * cols + sc];
            } else {
                v[k][j] = rec[(long long)rr * cols + cc];
            }
        }
#pragma unroll
    for (int k = 0; k < ROWS; ++k)
#pragma unroll
        for (int j = 0; j < PASSES; ++j) {
            const int rr = min(wave + WAVES * k, NROWS - 1), cc = min(lane + 64 * j, PITCH - 1);
            gv_lds[rr * PITCH + cc] = v[k][j];
        }
    __syncthreads();
}

// SRC: 0 global records, the pixel at least second - 1 from the raster edge; 1 global records
// near the edge (off-raster steps repeat the last sample); 2 records staged in LDS by stage_tile.
template <int SRC>
__device__ __forceinline__ GvfOut ground_view(const DomainG &G, const PixelAt &at, const v2f me,
                                              const float building, const AS4 double *H,
                                              const AS4 int32_t *gofs, const AS4 double *gbounds,
                                              int first, int second) {
    constexpr bool CHECK = SRC == 1;
    const int hh = second - 1, pitch = at.tx + 2 * hh;
    const int lbase = SRC == 2 ? (at.r - at.tr0 + hh) * pitch + (at.c - at.tc0 + hh) : 0;
    const int rows = (int)G.rows, cols = (int)G.cols;
    const long long P = G.pixels;
    const AS1 v2f *rec = G.rec;
    (void)gbounds;
    const int stride = (second + CHUNK - 1) / CHUNK * CHUNK;   // padded step axis of gofs
    GvSums gv(first, second);
    for (int dir = 0; dir < 18; ++dir) {
        const AS4 int32_t *off = gofs + dir * stride * 2;
        const uint32_t lb = (G.gvf_len[(long long)(dir >> 2) * P + at.p] >> (8 * (dir & 3))) & 0xffu;
        const int L = (int)(lb & 0x7fu);
        const float Lsteps = (float)L;
        float lup = 0.f, albsh = 0.f, f_lup = 0.f, f_albsh = 0.f;
        uint32_t kw = 255u;                           // first step on a sunlit wall (255: none)
        // near field when first > 1: the sums over the first ``first`` steps (step 0 is the
        // receiver itself); with first == 1 GvSums::finish adds it
        if (first > 1) {
            long long qn = at.p;
            for (int st = 0; st < first; ++st) {
                const int rr = at.r + off[2 * st], cc = at.c + off[2 * st + 1];
                if (!CHECK || ((unsigned)rr < (unsigned)rows && (unsigned)cc < (unsigned)cols))
                    qn = (long long)rr * cols + cc;
                const v2f r = st == 0 ? me
                                      : (SRC == 2 ? gv_lds[lbase + off[2 * st] * pitch + off[2 * st + 1]] : rec[qn]);
                const float on = __builtin_amdgcn_fmed3f(Lsteps - (float)st, 0.0f, 1.0f);
                f_lup = fmaf(on, r.x, f_lup);
                f_albsh = fmaf(on, fabsf(r.y), f_albsh);
            }
        }
        long long qprev = at.p;
        for (int s0 = 0; s0 < second; s0 += CHUNK) {
            if (__ballot(s0 < L) == 0) break;         // every lane's walk has ended
            v2f buf[CHUNK];
#pragma unroll
            for (int i = 0; i < CHUNK; ++i) {
                const int dr = off[2 * (s0 + i)], dc = off[2 * (s0 + i) + 1];
                if (SRC == 2) {
                    buf[i] = gv_lds[lbase + dr * pitch + dc];
                } else {
                    long long q;
                    if (!CHECK) {
                        q = at.p + (long long)(dr * cols + dc);
                    } else {
                        const int rr = at.r + dr, cc = at.c + dc;
                        const bool in = (unsigned)rr < (unsigned)rows && (unsigned)cc < (unsigned)cols;
                        q = in ? (long long)rr * cols + cc : qprev;
                        qprev = q;
                    }
                    buf[i] = rec[q];
                }
            }
            // step s is summed while s < L: on = clamp(L - s, 0, 1) (steps past ``second`` are
            // padding with L <= second, so they are off); the sunlit-wall bits of the chunk are
            // kept unmasked, a wall past L is discarded below by kw < L (kw is the first sunlit
            // wall of the whole walk, so kw < L exactly when one lies among the alive steps)
            const float rem = Lsteps - (float)s0;
            uint32_t cm = 0u;
#pragma unroll
            for (int i = 0; i < CHUNK; ++i) {
                const float on = __builtin_amdgcn_fmed3f(rem - (float)i, 0.0f, 1.0f);
                lup = fmaf(on, buf[i].x, lup);
                albsh = fmaf(on, fabsf(buf[i].y), albsh);
                // shift the record's sign bit in from the right: bit 7 - i is step s0 + i
                cm = __builtin_amdgcn_alignbit(cm, __float_as_uint(buf[i].y), 31);
            }
            if (kw == 255u && cm != 0u) kw = (uint32_t)(s0 + __builtin_clz(cm) - (32 - CHUNK));
        }
        if (first > 1) gv.add_near(dir, L, kw, first, f_lup, f_albsh);
        gv.add_far(dir, L, lb, kw, lup, albsh);
    }
    return gv.finish(G, at, H, me, building, first == 1);
}

// The ground-view shifts of the 18 directions for the first GV_FIXED_STEPS steps
// (sunonsurface_2018a, umep_solweig.py:1477-1492), the prefix of every scale's table.
constexpr int GV_FIXED_STEPS = 22;
constexpr int8_t GV_STEPS[18][GV_FIXED_STEPS][2] = {
#include "gvf_steps.inc"
};

// ground_view for a tile of width TX staged in LDS, first = 1 and second = GV_FIXED_STEPS
// (scale 1): the same sums as ground_view<2>, with every step's LDS address an immediate offset
// from the pixel's position in the staged tile (no per-step address arithmetic or table loads).
// EDGE (a tile whose halo leaves the raster): the steps past the pixel's last in-raster step (the
// static gvf_edge plane, E) repeat that step's sample, as ground_view<1> and the reference's
// shifted windows do, in the same order of operations.
template <int TX, bool EDGE>
__device__ __forceinline__ GvfOut ground_view_fixed(const DomainG &G, const PixelAt &at, const v2f me,
                                                    const float building, const AS4 double *H) {
    constexpr int NS = GV_FIXED_STEPS, HH = NS - 1, PITCH = TX + 2 * HH;
    const long long P = G.pixels;
    // the pixel's record sits at (HH, HH) from this base; every step's index is base + constant
    const int base = (at.r - at.tr0) * PITCH + (at.c - at.tc0);
    uint32_t lw[5], ew[5];                            // the 18 direction bytes, loaded at once
#pragma unroll
    for (int w = 0; w < 5; ++w) {
        lw[w] = G.gvf_len[w * P + at.p];
        ew[w] = EDGE ? G.gvf_edge[w * P + at.p] : 0u;
    }
    GvSums gv(1, NS);
#pragma unroll
    for (int dir = 0; dir < 18; ++dir) {
        const uint32_t lb = (lw[dir >> 2] >> (8 * (dir & 3))) & 0xffu;
        const int L = (int)(lb & 0x7fu);
        const int E = EDGE ? (int)((ew[dir >> 2] >> (8 * (dir & 3))) & 0xffu) : NS;
        // kept opaque: as an integer-valued float the compiler rewrites the clamps below into
        // integer subtract, sign-extend and convert sequences
        float Lsteps = (float)L;
        asm volatile("" : "+v"(Lsteps));
        float lup = 0.f, albsh = 0.f;
        v2f last = me;                                // EDGE: the last in-raster sample so far
        // the sign bits (sunlit walls) of every executed step, step 0 highest; nexec steps
        uint32_t cm = 0u;
        int nexec = 0;
#pragma unroll
        for (int s0 = 0; s0 < NS; s0 += CHUNK) {
            if (__ballot(s0 < L) == 0) break;
            const int n = NS - s0 < CHUNK ? NS - s0 : CHUNK;
            v2f buf[CHUNK];
            // volatile: one ds_read_b64 per step with the step's offset as its immediate from the
            // pixel's base (paired reads would need an address add per pair: their 8-bit offsets
            // do not reach rows 21 pixels away)
            const AS3 volatile v2f *lds = (const AS3 volatile v2f *)(AS3 void *)gv_lds;
#pragma unroll
            for (int i = 0; i < n; ++i) {
                const int k = (GV_STEPS[dir][s0 + i][0] + HH) * PITCH + GV_STEPS[dir][s0 + i][1] + HH;
                buf[i] = lds[base + k];
            }
            const float rem = Lsteps - (float)s0;
#pragma unroll
            for (int i = 0; i < n; ++i) {
                if (EDGE) {
                    // last = s0 + i < E ? buf[i] : last, as a bit select (a plain conditional
                    // becomes an exec-mask branch per step)
                    const uint32_t keep = s0 + i < E ? 0xffffffffu : 0u;
                    last.x = __uint_as_float((__float_as_uint(buf[i].x) & keep) | (__float_as_uint(last.x) & ~keep));
                    last.y = __uint_as_float((__float_as_uint(buf[i].y) & keep) | (__float_as_uint(last.y) & ~keep));
                    buf[i] = last;
                }
                const float on = __builtin_amdgcn_fmed3f(rem - (float)i, 0.0f, 1.0f);
                lup = fmaf(on, buf[i].x, lup);
                albsh = fmaf(on, fabsf(buf[i].y), albsh);
                cm = __builtin_amdgcn_alignbit(cm, __float_as_uint(buf[i].y), 31);
            }
            nexec = s0 + n;
        }
        // first sunlit-wall step; with none the unsigned difference wraps far above L
        const uint32_t lead = cm == 0u ? 0xffffffffu : (uint32_t)__builtin_clz(cm);
        const uint32_t kw = lead - (uint32_t)(32 - nexec);
        gv.add_far(dir, L, lb, kw, lup, albsh);
        // the half sums are complete here: without this the compiler keeps every direction's
        // partial terms live to the end of the walk (150 VGPRs instead of about 80)
        asm volatile("" : "+v"(gv.hB[0]), "+v"(gv.hB[1]), "+v"(gv.hB[2]), "+v"(gv.hB[3]),
                     "+v"(gv.hL[0]), "+v"(gv.hL[1]), "+v"(gv.hL[2]), "+v"(gv.hL[3]),
                     "+v"(gv.hA[0]), "+v"(gv.hA[1]), "+v"(gv.hA[2]), "+v"(gv.hA[3]));
    }
    return gv.finish(G, at, H, me, building, true);
}

// The 16-byte view of a pixel's hour record used by the adjoint kernels: Lup field, albedo,
// albedo x shadow, and the flag word (byte 0 ground, byte 1 sunlit wall, byte 2 shadow class).
__device__ __forceinline__ v4f record16(const DomainG &G, long long q) {
    const v2f r = G.rec[q];
    v4f o;
    o.x = r.x;
    o.y = G.alb[q];
    o.z = fabsf(r.y);
    o.w = __uint_as_float((G.ground[q] ? 1u : 0u) | ((__float_as_uint(r.y) >> 31) << 8) |
                          (class_of(G.masks[q]) << 16));
    return o;
}

// Static ground-view planes of a domain: per direction the steps before the first non-ground
// sample, and the unshadowed-albedo sums of the walk (gvfalbnosh) as a + albedo_b b.
__global__ __launch_bounds__(256) void k_gvf_static(const uint8_t *__restrict__ ground,
                                                    const float *__restrict__ alb,
                                                    const float *__restrict__ walls,
                                                    const float *__restrict__ aspect_deg, int rows, int cols,
                                                    const int32_t *__restrict__ gofs,
                                                    const double *__restrict__ gbounds, int first, int second,
                                                    uint32_t *__restrict__ len, float *__restrict__ nosh,
                                                    uint32_t *__restrict__ edge) {
    const long long P = (long long)rows * cols;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int r = (int)(p / cols), c = (int)(p % cols);
    const float building = ground[p] ? 1.0f : 0.0f;
    const float local = alb[p] * (1.0f - building);
    const bool wallbol = walls[p] > 0.0f;
    const double aspect = ((double)aspect_deg[p] * 3.141592653589793) / 180.0;
    const float firstf = (float)first, secondf = (float)second;
    const float inv_f = 1.0f / firstf, inv_f1 = 1.0f / (firstf + 1.0f), inv_s = 1.0f / secondf;
    const int stride = (second + CHUNK - 1) / CHUNK * CHUNK;
    float a[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, b[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    uint32_t words[5] = {0u, 0u, 0u, 0u, 0u}, ewords[5] = {0u, 0u, 0u, 0u, 0u};
    for (int dir = 0; dir < 18; ++dir) {
        const int32_t *off = gofs + dir * stride * 2;
        float af = building, wa = 0.f, fa = 0.f;
        int L = 0, fL = 0, E = second;
        bool snapped = false;
        long long q = p;
        for (int st = 0; st < second; ++st) {
            const int rr = r + off[2 * st], cc = c + off[2 * st + 1];
            if ((unsigned)rr < (unsigned)rows && (unsigned)cc < (unsigned)cols) q = (long long)rr * cols + cc;
            else if (E == second) E = st;
            af *= ground[q] ? 1.0f : 0.0f;
            if (af != 0.f) { ++L; wa += alb[q]; }
            if (st == first - 1) { fa = wa; fL = L; snapped = true; }
        }
        if (!snapped) { fa = wa; fL = L; }
        const float cnt_any = secondf - (float)L, f_cnt_any = firstf - (float)fL;
        const float a_near = f_cnt_any > 0.f ? fa * inv_f1 : fa * inv_f;
        const float b_near = f_cnt_any > 0.f ? f_cnt_any * inv_f1 : 0.f;
        const float a_far = wa * inv_s, b_far = cnt_any > 0.f ? cnt_any * inv_s : 0.f;
        const float ad = distance_weight(a_near, a_far) * building + local;
        const float bd = distance_weight(b_near, b_far) * building;
        // facesh of the direction (umep_solweig.py:1554-1563): the wall faces away from it
        const double *gbd = gbounds + dir * 3;
        int facesh;
        if (gbd[2] == 0.0)
            facesh = (int)(aspect < gbd[0] || aspect >= gbd[1]) - (wallbol ? 1 : 0) + 1;
        else
            facesh = 1 - (int)(aspect > gbd[0] || aspect <= gbd[1]);
        words[dir >> 2] |= (uint32_t)(L | (facesh == 0 ? 0x80 : 0)) << (8 * (dir & 3));
        ewords[dir >> 2] |= (uint32_t)E << (8 * (dir & 3));
        a[0] += ad; b[0] += bd;
        const int az = 5 + 20 * dir;
        if (az < 180) { a[1] += ad; b[1] += bd; }
        if (az >= 90 && az < 270) { a[2] += ad; b[2] += bd; }
        if (az >= 180) { a[3] += ad; b[3] += bd; }
        if (az >= 270 || az < 90) { a[4] += ad; b[4] += bd; }
    }
    for (int g = 0; g < 5; ++g) {
        nosh[g * P + p] = a[g];
        nosh[(5 + g) * P + p] = b[g];
        len[g * P + p] = words[g];
        if (edge) edge[g * P + p] = ewords[g];
    }
}

// Cylinder wedge: fraction of sunlit walls seen by a standing cylinder (umep_solweig.py:383-426),
// in FP32. It depends on y = tan(svfalfa) tan(zen) only: with v = 2 / y the reference's terms
// reduce to F = 1 - (v + (1 - v) area) / pi, the wall term present only for v > 1 (y < 2), where
// area = (sin phi - phi cos phi) / (1 - cos phi), phi = atan(t), t = sqrt(v^2 - 1/4). With
// R = sqrt(v^2 + 3/4) = 1 / cos phi, R - t = 1 / (R + t) and t - v = -(1/4) / (t + v), so
// F = 1 - [v / (R + t) - (1/4) / (t + v) + phi (v - 1)] / (pi (R - 1)),
// free of the cancellation the direct form has for large v; beyond v = 2^20 the asymptote
// 1/2 + 1 / (2 pi v) holds to 1e-12. y = 0 (open sky, svfalfa = 0) is the reference's NaN, which
// model.py:167 replaces by 0.5; the limit of F for y -> 0 is 0.5 as well.
__device__ __forceinline__ float wedge(double ta, double tb) {
    const float y = (float)(ta * tb);
    if (!(y > 0.0f)) return 0.5f;
    constexpr float INV_PI = 0.318309886183790672f;
    const float v = 2.0f * __builtin_amdgcn_rcpf(y);
    if (v <= 1.0f) return 1.0f - v * INV_PI;
    if (v > 1048576.0f) return 0.5f + (0.5f * INV_PI) * __builtin_amdgcn_rcpf(v);
    const float v2 = v * v;
    // hardware square roots (v_sqrt_f32, within 1 ulp; both arguments are >= 0.75)
    const float R = __builtin_amdgcn_sqrtf(v2 + 0.75f), t = __builtin_amdgcn_sqrtf(v2 - 0.25f);
    const float phi = atanf(t);
    const float num = v * __builtin_amdgcn_rcpf(R + t) - 0.25f * __builtin_amdgcn_rcpf(t + v) + phi * (v - 1.0f);
    return 1.0f - num * INV_PI * __builtin_amdgcn_rcpf(R - 1.0f);
}

// Tmrt from the absorbed radiation (umep_solweig.py:360): the fourth root as two hardware square
// roots (v_sqrt_f32, within 1 ulp each; the correctly rounded sqrtf takes about 15 instructions
// more per root)
__device__ __forceinline__ float tmrt_of(float Sstr, float absL) {
    return __builtin_amdgcn_sqrtf(__builtin_amdgcn_sqrtf(Sstr / (absL * (float)SBC))) - 273.2f;
}

__device__ __forceinline__ float lag(float current, float &state, double w, bool first,
                                     bool update) {
    const double cur = (double)current;
    const double prev = first ? cur : (double)state;
    const double delayed = cur * (1.0 - w) + prev * w;
    state = (float)(update ? delayed : prev);
    return (float)delayed;
}

// Patches per altitude band (create_patches, umep_solweig.py:1737-1786): band b holds patches
// [BAND_START[b], BAND_START[b+1]); the steradian-weighted trig terms are equal within a band.
constexpr int BAND_START[9] = {0, 31, 61, 89, 113, 132, 145, 152, 153};

// Daytime patch sums of one pixel (shortwave.cu and longwave.cu day terms, umep_solweig.py:666-911
// and 946-1356). Diffuse: sum of diffsh_k x (luminance, lumChi cos omega) over the patches, with
// diffsh 1 for sky, 1 - (1 - 0.03f) for vegetation-only and 0 for building patches. Building
// patches are split into sunlit and shaded walls by the exact thresholds of the hour's table;
// since the trig weights are per band, those sums are band counts times the band weight.
// DIFF false: the diffuse sums come from sw_day_diffuse; only the wall patches are summed.
template <bool DIFF>
__device__ __forceinline__ void day_patch_sums(const AS1 uint32_t *vis, long long P, long long p,
                                               const AS4 float *pf, const AS4 double *pd,
                                               const AS4 double *H, double tan_asvf, float dveg, float &a_lv, float &a_kd,
                                               float &a_swsun, float &a_swsh, float &a_lsunc,
                                               float &a_lsuns, float &a_lshc, float &a_lshs) {
#pragma unroll
    for (int w = 0; w < SW_VIS_WORDS; ++w) {
        const uint32_t vsh = vis[(0 * SW_VIS_WORDS + w) * P + p];
        const uint32_t vveg = vis[(1 * SW_VIS_WORDS + w) * P + p];
        const uint32_t vvb = vis[(2 * SW_VIS_WORDS + w) * P + p];
        const uint32_t sky = vsh & vveg, vegonly = vsh & ~vveg, bld = ~vsh & vvb;
        const int k0 = w * 32, n = SW_PATCHES - k0 < 32 ? SW_PATCHES - k0 : 32;
        // rolled loops: this path (no search tables) is the ABI's fallback; unrolled, its per-bit
        // lane masks fill the scalar registers of every kernel that inlines it
#pragma nounroll
        for (int j = 0; j < (DIFF ? n : 0); ++j) {
            const AS4 float *c = pf + (k0 + j) * SW_PATCH_F;
            const float ds = ((sky >> j) & 1u) ? 1.0f : (((vegonly >> j) & 1u) ? dveg : 0.0f);
            a_lv = fmaf(ds, c[SWP_LV], a_lv);
            a_kd = fmaf(ds, c[SWP_KD], a_kd);
        }
        if (__ballot(bld != 0u) == 0) continue;       // no building patch in this word
        const uint32_t elig = (uint32_t)H[SWH_ELIG + w];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int lo = BAND_START[b] > k0 ? BAND_START[b] : k0;
            const int hi = BAND_START[b + 1] < k0 + n ? BAND_START[b + 1] : k0 + n;
            if (lo >= hi) continue;
            // the two thresholds depend on the patch altitude only: one pair per band
            const double t_sun = pd[SWP_SUN_BELOW * SW_PATCHES + BAND_START[b]];
            const double t_sh = pd[SWP_SHADE_FROM * SW_PATCHES + BAND_START[b]];
            uint32_t sun = 0u, shd = 0u;
#pragma nounroll
            for (int k = lo; k < hi; ++k) {
                const double x = tan_asvf + pd[SWP_YI * SW_PATCHES + k];
                sun |= (x < t_sun ? 1u : 0u) << (k - k0);
                shd |= (x >= t_sh ? 1u : 0u) << (k - k0);
            }
            const uint32_t span = (uint32_t)(hi - lo);
            const uint32_t mask = (span == 32u ? 0xffffffffu : ((1u << span) - 1u)) << (lo - k0);
            const uint32_t bm = bld & mask;
            const float wc = pf[BAND_START[b] * SW_PATCH_F + SWP_WCOS];
            const float wsn = pf[BAND_START[b] * SW_PATCH_F + SWP_WSIN];
            const float c_sun = (float)__popc(bm & sun), c_sh = (float)__popc(bm & shd);
            const float c_lsun = (float)__popc(bm & elig & sun);
            const float c_lsh = (float)__popc(bm & (~elig | shd));
            a_swsun = fmaf(wc, c_sun, a_swsun);
            a_swsh = fmaf(wc, c_sh, a_swsh);
            a_lsunc = fmaf(wc, c_lsun, a_lsunc);
            a_lsuns = fmaf(wsn, c_lsun, a_lsuns);
            a_lshc = fmaf(wc, c_lsh, a_lshc);
            a_lshs = fmaf(wsn, c_lsh, a_lshs);
        }
    }
}

// Bisection steps over the patches of each altitude band (ceil(log2(count + 1))).
constexpr int BAND_STEPS[8] = {5, 5, 5, 5, 5, 4, 3, 1};


// Copy one hour row of the patch search tables, from word ``from`` (a multiple of 4) on, into
// LDS at the same word positions (every thread of the workgroup calls; the caller has
// synchronised the workgroup since the LDS was last read).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void stage_search(const AS1 uint32_t *src, int tid, int from) {
    AS3 u32x4 *dst = (AS3 u32x4 *)(AS3 void *)gv_lds;
    const AS1 u32x4 *s4 = (const AS1 u32x4 *)src;
    for (int i = tid + from / 4; i < SW_SEARCH_WORDS / 4; i += BLOCK) dst[i] = s4[i];
    __syncthreads();
}

// day_patch_sums with the hour's search tables in LDS (stage_search): the diffuse sums are 8
// nibble-table reads per visibility word and mask, and per altitude band the sunlit and the
// shaded wall patches are the patches below a bisected position in the band's yi order (the
// exact FP64 tests tan(asvf) + yi < threshold, as comparisons of tan(asvf) with the tables'
// breakpoints, evaluated at the bisection probes only). DIFF false: the diffuse sums come from
// sw_day_diffuse and the nibble tables are neither staged nor read.
template <bool DIFF>
__device__ __forceinline__ void day_patch_sums_tab(const AS1 uint32_t *vis, long long P, long long p,
                                                   const AS4 float *pf, const AS4 double *pd,
                                                   double tan_asvf, float dveg, float &a_lv, float &a_kd,
                                                   float &a_swsun, float &a_swsh, float &a_lsunc,
                                                   float &a_lsuns, float &a_lshc, float &a_lshs) {
    const AS3 uint32_t *T = (const AS3 uint32_t *)(AS3 void *)gv_lds;
    const AS3 double *cs = (const AS3 double *)(T + SWS_CSUN), *ch = (const AS3 double *)(T + SWS_CSH);
    const AS3 uint32_t *mk = T + SWS_MASK;
    const AS3 uint32_t *el = T + SWS_ELIG;
    float sl = 0.f, sk = 0.f, vl = 0.f, vk = 0.f;
    uint32_t bld[SW_VIS_WORDS];
    // ``tb`` (always 0) orders each word's table reads after the previous word's sums: left free,
    // the compiler issues all 78 reads at once (over 200 VGPRs)
    int tb = 0;
    // the nibble tables through one base register: a read is then one bit-field extract (kept as
    // such: folded into the address, the extract becomes a shift and a mask) and one shift-add,
    // with the table's position as the read's immediate offset
    uint32_t nib_base = (uint32_t)(uintptr_t)(T + SWS_NIB);
    asm volatile("" : "+v"(nib_base));
    const AS3 v2f *nib = (const AS3 v2f *)(AS3 void *)(uintptr_t)nib_base;
#pragma unroll
    for (int w = 0; w < SW_VIS_WORDS; ++w) {
        const uint32_t vsh = vis[(0 * SW_VIS_WORDS + w) * P + p];
        const uint32_t vvb = vis[(2 * SW_VIS_WORDS + w) * P + p];
        bld[w] = ~vsh & vvb;
        if (!DIFF) continue;
        const uint32_t vveg = vis[(1 * SW_VIS_WORDS + w) * P + p];
        const uint32_t sky = vsh & vveg, vegonly = vsh & ~vveg;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = 8 * w + q;
            if (i >= 39) break;
            uint32_t n1, n2;
            asm("v_bfe_u32 %0, %1, %2, 4" : "=v"(n1) : "v"(sky), "s"(4 * q));
            asm("v_bfe_u32 %0, %1, %2, 4" : "=v"(n2) : "v"(vegonly), "s"(4 * q));
            const v2f t1 = nib[tb + i * 16 + (int)n1];
            const v2f t2 = nib[tb + i * 16 + (int)n2];
            sl += t1.x; sk += t1.y; vl += t2.x; vk += t2.y;
        }
        asm volatile("" : "+v"(tb), "+v"(sl), "+v"(sk), "+v"(vl), "+v"(vk));
    }
    if (DIFF) {
        a_lv = fmaf(dveg, vl, sl);
        a_kd = fmaf(dveg, vk, sk);
    }
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const int lo = BAND_START[b], n = BAND_START[b + 1] - BAND_START[b];
        const int wi = lo >> 5, sh = lo & 31;
        const uint32_t full = n == 32 ? 0xffffffffu : ((1u << n) - 1u);
        const uint32_t bm = (sh + n <= 32 ? (bld[wi] >> sh) : __builtin_amdgcn_alignbit(bld[wi + 1], bld[wi], sh)) & full;
        if (__ballot(bm != 0u) == 0) continue;       // no building patch of the band in this wave
        // tan(asvf) + yi_j < threshold exactly when tan(asvf) < c_j (the tables' breakpoints, non-
        // increasing in j): j counts the band's sunlit (not shaded) patches in yi order
        int js = tb, jh = tb;
#pragma unroll
        for (int s = 1 << (BAND_STEPS[b] - 1); s > 0; s >>= 1) {
            js += (tan_asvf < cs[b * 32 + js + s - 1]) ? s : 0;
            jh += (tan_asvf < ch[b * 32 + jh + s - 1]) ? s : 0;
        }
        const uint32_t sun = mk[b * 33 + js], shd = full & ~mk[b * 33 + jh], elig = el[b];
        const float wc = pf[lo * SW_PATCH_F + SWP_WCOS];
        const float wsn = pf[lo * SW_PATCH_F + SWP_WSIN];
        const float c_sun = (float)__popc(bm & sun), c_sh = (float)__popc(bm & shd);
        const float c_lsun = (float)__popc(bm & elig & sun);
        const float c_lsh = (float)__popc(bm & (~elig | shd));
        a_swsun = fmaf(wc, c_sun, a_swsun);
        a_swsh = fmaf(wc, c_sh, a_swsh);
        a_lsunc = fmaf(wc, c_lsun, a_lsunc);
        a_lsuns = fmaf(wsn, c_lsun, a_lsuns);
        a_lshc = fmaf(wc, c_lsh, a_lshc);
        a_lshs = fmaf(wsn, c_lsh, a_lshs);
        asm volatile("" : "+v"(tb), "+v"(a_swsun), "+v"(a_lshs));
    }
}

// One hour of the pixel pass for one pixel of a workgroup's tile (every lane of the workgroup
// calls it: it holds workgroup barriers). The pixel's lag state (the five Lup delays, and in FULL
// mode TgOut's) is st; day hours update it in place, the caller loads and stores it. G.rec and
// G.masks address the hour's shadow-pass record and mask planes.
template <bool FULL>
__device__ __forceinline__ void pixel_hour(const DomainG &G, const PixelAt &at, const bool valid, const int hour,
                                           const AS4 int32_t *gofs, const AS4 double *gbounds, int first,
                                           int second, int lds_cap, float (&st)[6]) {
    const AS4 double *H = G.hours + (long long)hour * SW_HOUR_STRIDE;
    const long long P = G.pixels, p = at.p;
    const bool day = H[SWH_DAY] != 0.0;
    const float emis = G.emis[p];
    const float radI = (float)H[SWH_RADI], radD = (float)H[SWH_RADD], radG = (float)H[SWH_RADG];
    const float albedo_b = (float)H[SWH_ALBEDO_B];
    const float flux4 = (float)H[SWH_FLUX4];
    float Kdown = 0.f, Kup = 0.f, KupE = 0.f, KupS = 0.f, KupW = 0.f, KupN = 0.f;
    float KsideI = 0.f, Lup, LupE, LupS, LupW, LupN, shadow = 0.f, Fsh = 0.f, tgout = 0.f, tg = 0.f;
    if (day) {
        const v2f me = G.rec[p];
        const float building = G.ground[p] ? 1.0f : 0.0f;
        shadow = shadow_of_class(class_of(G.masks[p]), (float)(1.0 - (1.0 - H[SWH_PSI])));
        const int m = second - 1;
        // tiles of width 16 or 32 stage their records and halo in LDS: with the halo inside the
        // raster always, with the scale-1 geometry also at the raster edge (clamped, gvf_edge)
        const bool fits = lds_cap > 0 && at.tx <= 32 && (at.tx + 2 * m) * (at.ty + 2 * m) <= lds_cap;
        const bool inner = at.tr0 >= m && at.tc0 >= m && at.tr0 + at.ty + m <= (int)G.rows &&
                           at.tc0 + at.tx + m <= (int)G.cols;
        const bool fixed = first == 1 && second == GV_FIXED_STEPS && (at.tx == 16 || at.tx == 32) &&
                           (inner || G.gvf_edge != nullptr);
        GvfOut g;
        if (fits && fixed) {
            constexpr int HH = GV_FIXED_STEPS - 1;
            if (at.tx == 16 && inner) {
                stage_tile_fixed<16, HH, false>(G, at);
                g = ground_view_fixed<16, false>(G, at, me, building, H);
            } else if (at.tx == 16) {
                stage_tile_fixed<16, HH, true>(G, at);
                g = ground_view_fixed<16, true>(G, at, me, building, H);
            } else if (inner) {
                stage_tile_fixed<32, HH, false>(G, at);
                g = ground_view_fixed<32, false>(G, at, me, building, H);
            } else {
                stage_tile_fixed<32, HH, true>(G, at);
                g = ground_view_fixed<32, true>(G, at, me, building, H);
            }
        } else if (fits && inner) {
            stage_tile(G, at, m, at.tx + 2 * m);
            g = ground_view<2>(G, at, me, building, H, gofs, gbounds, first, second);
        } else {
            const bool interior = at.r >= m && at.c >= m && at.r < (int)G.rows - m && at.c < (int)G.cols - m;
            g = interior ? ground_view<0>(G, at, me, building, H, gofs, gbounds, first, second)
                         : ground_view<1>(G, at, me, building, H, gofs, gbounds, first, second);
        }
        const float amb = flux4 * emis;
        const double w = H[SWH_LAG_WEIGHT];
        const bool lf = H[SWH_LAG_FIRST] != 0.0, lu = H[SWH_LAG_UPDATE] != 0.0;
        Lup = lag(g.all.lup * INV18 + amb, st[0], w, lf, lu);
        LupE = lag(g.e.lup * INV9 + amb, st[1], w, lf, lu);
        LupS = lag(g.s.lup * INV9 + amb, st[2], w, lf, lu);
        LupW = lag(g.w.lup * INV9 + amb, st[3], w, lf, lu);
        LupN = lag(g.n.lup * INV9 + amb, st[4], w, lf, lu);
        if (FULL) {
            const double amp = (double)G.tgk[p] * H[SWH_ALTMAX] + (double)G.tstart[p];
            double t = amp * tg_sine(H, G.tmaxlst[p]) * H[SWH_CITGG];
            if (G.landcover == 1 && t < 0.0) t = 0.0;
            tg = (float)t;
            tgout = lag((float)(t * (double)shadow + H[SWH_TA]), st[5], w, lf, lu);
        }
        Fsh = wedge(G.tan_svfalfa[p], H[SWH_TANZEN]);
        const float svfbuveg = G.svfbuveg[p];
        const float sinalt = (float)H[SWH_SINALT];
        const float around = albedo_b * (1.0f - svfbuveg) * (radG * (1.0f - Fsh) + radD * Fsh);
        const float refl = radD * svfbuveg + around;
        Kup = (g.all.alb * INV18) * radI * sinalt + refl * (g.all.nosh * INV18);
        KupE = (g.e.alb * INV9) * radI * sinalt + refl * (g.e.nosh * INV9);
        KupS = (g.s.alb * INV9) * radI * sinalt + refl * (g.s.nosh * INV9);
        KupW = (g.w.alb * INV9) * radI * sinalt + refl * (g.w.nosh * INV9);
        KupN = (g.n.alb * INV9) * radI * sinalt + refl * (g.n.nosh * INV9);
        Kdown = radI * shadow * sinalt + around;  // + dRad below
        KsideI = shadow * radI * (float)H[SWH_COSALT];
    } else {
        Lup = flux4 * emis;
        LupE = LupS = LupW = LupN = Lup;
        if (FULL) tgout = (float)H[SWH_TA];
    }
    // ---- anisotropic patch sums over the 153 sky patches --------------------------------
    const int row = uniform_index(H[SWH_PATCH_TABLE]);
    const AS4 float *pf = G.ptab_f + (long long)row * SW_PATCHES * SW_PATCH_F;
    const AS4 double *pd = G.ptab_d + (long long)row * SW_PATCH_D * SW_PATCHES;
    const double tan_asvf = G.tan_asvf[p];
    const float dveg = (float)H[SWH_DIFF_VEG];
    float a_lv = 0.f, a_kd = 0.f, a_ldn = 0.f, a_lsd = 0.f, a_vc = 0.f, a_vs = 0.f;
    float a_swsun = 0.f, a_swsh = 0.f, a_lsunc = 0.f, a_lsuns = 0.f, a_lshc = 0.f, a_lshs = 0.f;
    float a_rc = 0.f, a_rs = 0.f;
    float cE = 0.f, cW = 0.f, cN = 0.f, cS = 0.f;       // FULL: cardinal sky+veg+walls
    float rE = 0.f, rW = 0.f, rN = 0.f, rS = 0.f;       // FULL: cardinal reflected weights
    const float lw_amb = (float)H[SWH_LW_AMBIENT], lw_sun = (float)H[SWH_LW_SUNLIT];
    const AS1 uint32_t *vis = G.vis;
    if (!FULL) {
        // hour-independent parts from the static sums; sky flux linear in the sky emissivity
        const AS1 float *ss = G.skysums;
        const float A = (float)H[SWH_SKY_A], B = (float)H[SWH_SKY_B];
        a_ldn = A * ss[SWS_SKY_SIN * P + p] - B * ss[SWS_SKY_GSIN * P + p];
        a_lsd = A * ss[SWS_SKY_COS * P + p] - B * ss[SWS_SKY_GCOS * P + p];
        a_vc = ss[SWS_VEG_COS * P + p];
        a_vs = ss[SWS_VEG_SIN * P + p];
        a_rc = ss[SWS_OBS_COS * P + p];
        a_rs = ss[SWS_OBS_SIN * P + p];
        if (!day) {
            a_lshc = ss[SWS_BLD_COS * P + p];
            a_lshs = ss[SWS_BLD_SIN * P + p];
        }
    }
    // the diffuse sums of the hour from sw_day_diffuse (a domain either has them or not)
    const bool pre = !FULL && day && G.diffuse != nullptr;
    if (pre) {
        const v2f dd = G.diffuse[(long long)hour * P + p];
        a_lv = dd.x;
        a_kd = dd.y;
    }
    if (!FULL && day && G.ptab_s) {
        __syncthreads();                              // every wave's walk is done with the LDS
        const AS1 uint32_t *srow = G.ptab_s + (long long)row * SW_SEARCH_WORDS;
        if (pre) {
            stage_search(srow, at.tid, SWS_CSUN);
            day_patch_sums_tab<false>(vis, P, p, pf, pd, tan_asvf, dveg, a_lv, a_kd, a_swsun, a_swsh,
                                      a_lsunc, a_lsuns, a_lshc, a_lshs);
        } else {
            stage_search(srow, at.tid, 0);
            day_patch_sums_tab<true>(vis, P, p, pf, pd, tan_asvf, dveg, a_lv, a_kd, a_swsun, a_swsh,
                                     a_lsunc, a_lsuns, a_lshc, a_lshs);
        }
    } else if (!FULL && day) {
        if (pre)
            day_patch_sums<false>(vis, P, p, pf, pd, H, tan_asvf, dveg, a_lv, a_kd, a_swsun, a_swsh, a_lsunc,
                                  a_lsuns, a_lshc, a_lshs);
        else
            day_patch_sums<true>(vis, P, p, pf, pd, H, tan_asvf, dveg, a_lv, a_kd, a_swsun, a_swsh, a_lsunc,
                                 a_lsuns, a_lshc, a_lshs);
    }
    if (FULL) {
        float lv_sky = 0.f, lv_veg = 0.f, kd_sky = 0.f, kd_veg = 0.f;
        for (int w = 0; w < SW_VIS_WORDS; ++w) {
            const uint32_t vsh = vis[(0 * SW_VIS_WORDS + w) * P + p];
            const uint32_t vveg = vis[(1 * SW_VIS_WORDS + w) * P + p];
            const uint32_t vvb = vis[(2 * SW_VIS_WORDS + w) * P + p];
            const uint32_t sky = vsh & vveg, vegonly = vsh & ~vveg;
            const uint32_t vegf = ~vveg | ~vvb, bld = ~vsh & vvb, obst = ~vsh | ~vveg | ~vvb;
            const int k0 = w * 32, k1 = min(SW_PATCHES, k0 + 32);
            for (int k = k0; k < k1; ++k) {
                const int j = k - k0;
                const AS4 float *c = pf + k * SW_PATCH_F;
                const float wcos = c[SWP_WCOS], wsin = c[SWP_WSIN];
                if ((sky >> j) & 1) { lv_sky += c[SWP_LV]; kd_sky += c[SWP_KD]; }
                if ((vegonly >> j) & 1) { lv_veg += c[SWP_LV]; kd_veg += c[SWP_KD]; }
                if (FULL) {
                    if ((sky >> j) & 1) {
                        a_ldn += c[SWP_LDN];
                        a_lsd += c[SWP_LSD];
                        cE += c[SWP_LSD] * c[SWP_DE]; cW += c[SWP_LSD] * c[SWP_DW];
                        cN += c[SWP_LSD] * c[SWP_DN]; cS += c[SWP_LSD] * c[SWP_DS];
                    }
                    if ((vegf >> j) & 1) {
                        a_vc += wcos;
                        a_vs += wsin;
                        const float v = lw_amb * wcos;
                        cE += v * c[SWP_DE]; cW += v * c[SWP_DW]; cN += v * c[SWP_DN]; cS += v * c[SWP_DS];
                    }
                    if ((obst >> j) & 1) {
                        a_rc += wcos;
                        a_rs += wsin;
                        rE += wcos * c[SWP_DE]; rW += wcos * c[SWP_DW];
                        rN += wcos * c[SWP_DN]; rS += wcos * c[SWP_DS];
                    }
                }
                if ((bld >> j) & 1) {
                    bool sunlit = false, shaded = true;
                    if (day) {
                        const double x = tan_asvf + pd[SWP_YI * SW_PATCHES + k];
                        sunlit = x < pd[SWP_SUN_BELOW * SW_PATCHES + k];
                        shaded = x >= pd[SWP_SHADE_FROM * SW_PATCHES + k];
                    }
                    if (sunlit) a_swsun += wcos;
                    if (shaded) a_swsh += wcos;
                    const bool elig = c[SWP_ELIGIBLE] != 0.0f;
                    const bool lsun = elig && sunlit, lsh = elig ? shaded : true;
                    if (lsun) { a_lsunc += wcos; a_lsuns += wsin; }
                    if (lsh) { a_lshc += wcos; a_lshs += wsin; }
                    if (FULL) {
                        const float v = (lsun ? lw_sun * wcos : 0.f) + (lsh ? lw_amb * wcos : 0.f);
                        cE += v * c[SWP_DE]; cW += v * c[SWP_DW]; cN += v * c[SWP_DN]; cS += v * c[SWP_DS];
                    }
                }
            }
        }
        a_lv = lv_sky + dveg * lv_veg;
        a_kd = kd_sky + dveg * kd_veg;
    }
    const float dRad = a_lv * radD;
    Kdown += day ? dRad : 0.f;
    const float sw_sun = (float)H[SWH_SW_SUNLIT], sw_sh = (float)H[SWH_SW_SHADED];
    const float KsideD = day ? a_kd : 0.f;
    const float Kside = day ? KsideI + KsideD + sw_sun * a_swsun + sw_sh * a_swsh + sw_sh * a_vc : 0.f;
    const float reflsurf = (a_ldn + Lup) * (float)H[SWH_LW_REFL];
    const float Ldown = a_ldn + lw_amb * a_vs + lw_amb * a_lshs + lw_sun * a_lsuns + reflsurf * a_rs;
    const float Lside = a_lsd + lw_amb * a_vc + lw_amb * a_lshc + lw_sun * a_lsunc + reflsurf * a_rc;
    const float absK = (float)H[SWH_ABSK], absL = (float)H[SWH_ABSL];
    const float Fup = (float)H[SWH_FUP], Fside = (float)H[SWH_FSIDE], Fcyl = (float)H[SWH_FCYL];
    const float Ksum = (KupN + KupE + KupS + KupW) * 0.5f;
    const float Lsum = (LupN + LupE + LupS + LupW) * 0.5f;
    const float Sstr = absK * (Kside * Fcyl + (Kdown + Kup) * Fup + Ksum * Fside) +
                       absL * ((Ldown + Lup) * Fup + Lside * Fcyl + Lsum * Fside);
    const float Tmrt = tmrt_of(Sstr, absL);
    // UTCI with the hour's scalar weather (utci.cu:500-523), NaN off the ground mask.
    float u;
    if (!G.ground[p]) {
        u = __builtin_nanf("");
    } else if (H[SWH_UTCI0 + 6] == -999.0 || Tmrt <= -999.0f) {
        u = -999.0f;
    } else {
        const double dT = (double)(Tmrt - (float)H[SWH_TA32]);
        double acc = H[SWH_UTCI0 + 6];
        for (int k = 5; k >= 0; --k) acc = acc * dT + H[SWH_UTCI0 + k];
        u = (float)acc;
    }
    if (!valid) return;
    if (G.tmrt) G.tmrt[(long long)hour * P + p] = Tmrt;
    if (G.utci) G.utci[(long long)hour * P + p] = u;
    if (!FULL && G.full) {
        AS1 float *o = G.full + (long long)hour * SW_EXTRA_PLANES * P;
        o[SWE_KDOWN * P + p] = Kdown;
        o[SWE_KUP * P + p] = Kup;
        o[SWE_LDOWN * P + p] = Ldown;
        o[SWE_LUP * P + p] = Lup;
        o[SWE_SHADOW * P + p] = shadow;
    }
    if (FULL && G.full) {
        AS1 float *o = G.full;
        const float refl = reflsurf;
        o[SWF_KDOWN * P + p] = Kdown;
        o[SWF_KUP * P + p] = Kup;
        o[SWF_LDOWN * P + p] = Ldown;
        o[SWF_LUP * P + p] = Lup;
        o[SWF_TG * P + p] = tg;
        o[SWF_KEAST * P + p] = KupE * 0.5f;
        o[SWF_KSOUTH * P + p] = KupS * 0.5f;
        o[SWF_KWEST * P + p] = KupW * 0.5f;
        o[SWF_KNORTH * P + p] = KupN * 0.5f;
        o[SWF_LEAST * P + p] = LupE * 0.5f + cE + refl * rE;
        o[SWF_LSOUTH * P + p] = LupS * 0.5f + cS + refl * rS;
        o[SWF_LWEST * P + p] = LupW * 0.5f + cW + refl * rW;
        o[SWF_LNORTH * P + p] = LupN * 0.5f + cN + refl * rN;
        o[SWF_KSIDEI * P + p] = KsideI;
        o[SWF_TGOUT * P + p] = tgout;
        o[SWF_LSIDE * P + p] = Lside;
        o[SWF_KSIDED * P + p] = KsideD;
        o[SWF_DRAD * P + p] = day ? dRad : 0.f;
        o[SWF_KSIDE * P + p] = Kside;
        o[SWF_SHADOW * P + p] = shadow;
        o[SWF_WALLSUN * P + p] = 0.f;
        o[SWF_FSH * P + p] = Fsh;
    }
}

// Lanes past the raster (partial tiles) compute a copy of the nearest pixel, so that every wave
// takes part in the workgroup's barriers, and write nothing.
__device__ __forceinline__ bool clamp_to_raster(const DomainG &G, PixelAt &at) {
    if (at.valid) return true;
    at.r = min(at.r, (int)G.rows - 1);
    at.c = min(at.c, (int)G.cols - 1);
    at.p = (long long)at.r * G.cols + at.c;
    return false;
}

// The pixel pass of one hour: the lag state round-trips the domain's lag planes (day hours only).
template <bool FULL>
__global__ __launch_bounds__(BLOCK) void k_hour_pixels(const sw_domain *__restrict__ doms,
                                                       const int32_t *__restrict__ bs, int nd,
                                                       int total, int hour,
                                                       const int32_t *__restrict__ gofs_,
                                                       const double *__restrict__ gbounds_,
                                                       int first, int second, int lds_cap) {
    DomainG G;
    PixelAt at = locate(doms, bs, nd, total, G);
    const bool valid = clamp_to_raster(G, at);
    const long long P = G.pixels, p = at.p;
    const bool day = G.hours[(long long)hour * SW_HOUR_STRIDE + SWH_DAY] != 0.0;
    constexpr int NS = FULL ? 6 : 5;
    AS1 float *lg = G.lag;
    float st[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (day)
        for (int k = 0; k < NS; ++k) st[k] = lg[k * P + p];
    pixel_hour<FULL>(G, at, valid, hour, (const AS4 int32_t *)gofs_, (const AS4 double *)gbounds_, first, second,
                     lds_cap, st);
    if (day && valid)
        for (int k = 0; k < NS; ++k) lg[k * P + p] = st[k];
}

// The pixel pass of hours h0 .. h1-1 in one launch: each workgroup runs its tile through the
// hours in order with the lag state in registers (loaded from and stored back to the lag planes
// once), reading each hour's records and mask bytes from slot ``hour`` of the per-hour planes that
// the shadow pass wrote with per_hour set. Tiles are independent, so no grid-wide barrier is needed.
__global__ __launch_bounds__(BLOCK) void k_day_pixels(const sw_domain *__restrict__ doms,
                                                      const int32_t *__restrict__ bs, int nd, int total,
                                                      int h0, int h1, const int32_t *__restrict__ gofs_,
                                                      const double *__restrict__ gbounds_, int first,
                                                      int second, int lds_cap) {
    float st[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    {
        DomainG G;
        PixelAt at = locate(doms, bs, nd, total, G);
        clamp_to_raster(G, at);
        for (int k = 0; k < 5; ++k) st[k] = G.lag[k * G.pixels + at.p];
    }
#pragma nounroll
    for (int hour = h0; hour < h1; ++hour) {
        if (hour > h0) __syncthreads();               // the previous hour's LDS reads are done
        // Each hour locates its tile and reads the domain record again, from the launch arguments
        // and the lane index made opaque to the loop optimiser, as if the launch began there:
        // nothing derived from them (the record, plane and staging addresses, table pointers) is
        // held in registers across the hours, where it would cost occupancy.
        const sw_domain *dh = doms;
        const int32_t *bh = bs, *go = gofs_;
        const double *gb = gbounds_;
        int fs = first, ss = second, cap = lds_cap, lane = (int)threadIdx.x;
        asm volatile("" : "+s"(dh), "+s"(bh), "+s"(go), "+s"(gb), "+s"(fs), "+s"(ss), "+s"(cap));
        asm volatile("" : "+v"(lane));
        DomainG G;
        PixelAt at = locate_lane(dh, bh, nd, total, G, lane);
        const bool valid = clamp_to_raster(G, at);
        G.rec += (long long)hour * G.pixels;
        G.masks += (long long)hour * G.pixels;
        pixel_hour<false>(G, at, valid, hour, (const AS4 int32_t *)go, (const AS4 double *)gb, fs, ss, cap, st);
    }
    DomainG G;
    PixelAt at = locate(doms, bs, nd, total, G);
    if (clamp_to_raster(G, at))
        for (int k = 0; k < 5; ++k) G.lag[k * G.pixels + at.p] = st[k];
}

// =====================================================================================
// Diffuse patch sums of a day on the matrix cores (sw_day_diffuse). Per pixel and hour the
// reference sums diffsh_k x luminance_k and diffsh_k x lumChi_k cos(alt_k) omega_k over the 153
// patches (umep_solweig.py:148-154, shortwave.cu:99-107 and 155-157), diffsh 1 on sky patches,
// the hour's vegetation diffsh on vegetation-only patches and 0 on building patches: over a day,
// the product of the [rows][320] table of the hour values (two rows per hour) with the
// [320][pixels] matrix of the pixels' sky | vegetation-only bits. The table is integer (int8
// digits of 27-bit fixed-point values with a power-of-two scale per row), so the int32 sums of
// v_mfma_i32_16x16x64_i8 are exact and each result is rounded once, to FP32: the sums do not
// depend on the launch geometry.
// Mapping: a workgroup takes the pixel passes' tile; wave w takes the pixels of the tile's lanes
// 64w .. 64w + 63 as four 16-pixel column blocks (block t: lanes 64w + 16t + c, c = 0..15).
// Lane c + 16g holds as the A operand row c of an 8-hour block of the table at positions
// 64s + 16g .. 64s + 16g + 15 of step s (s = 0..4; the table stores these fragments in lane
// order, so a wave's load is 1 KB in one piece), as the B operand the same positions of column
// c's pixel, and receives rows 4g .. 4g + 3 of that pixel (the 16 x 16 result layout).
// =====================================================================================
typedef int v4i __attribute__((ext_vector_type(4)));

// 16 bits as 16 int8 values 0 / 1, bit j in byte j: a nibble n becomes the word
// n * 0x204081 & 0x01010101 (its shifted copies n, n << 7, n << 14, n << 21 do not overlap)
__device__ __forceinline__ v4i bits_to_bytes(uint32_t b16) {
    v4i v;
    v.x = (int)(__umul24(b16 & 0xfu, 0x204081u) & 0x01010101u);
    v.y = (int)(__umul24((b16 >> 4) & 0xfu, 0x204081u) & 0x01010101u);
    v.z = (int)(__umul24((b16 >> 8) & 0xfu, 0x204081u) & 0x01010101u);
    v.w = (int)(__umul24((b16 >> 12) & 0xfu, 0x204081u) & 0x01010101u);
    return v;
}

__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void k_day_diffuse(
    const sw_domain *__restrict__ doms, const int32_t *__restrict__ bs, int nd, int total, int n_hours) {
    DomainG G;
    const PixelAt at = locate(doms, bs, nd, total, G);
    if (G.diffuse_tab == nullptr || G.diffuse == nullptr) return;   // one domain per workgroup
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int rows = (int)G.rows, cols = (int)G.cols, sh = G.tile_shift;
    const long long P = G.pixels;
    const int nrows = SW_DIFF_ROWS(n_hours);
    // per block t and step s this lane's 16 bits: word 2s + g / 2 of the pixel's sky (words
    // 0..4) then vegetation-only (5..9) masks, its half g % 2; steps 2j and 2j + 1 share a word
    uint32_t part[4][3];
    // block t's pixel of this lane (recomputed where it is stored: kept, the four 64-bit indices
    // would cost the registers of a fourth wave per SIMD)
    auto pixel_of = [&](int t, bool &in) {
        const int l = 64 * wave + 16 * t + c;
        const int r = at.tr0 + (l >> sh), cc = at.tc0 + (l & ((1 << sh) - 1));
        in = r < rows && cc < cols;
        return in ? (long long)r * cols + cc : 0LL;       // pixel 0 stands in outside the raster
    };
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        bool in;
        const long long q = pixel_of(t, in);
        const AS1 uint32_t *vq = G.vis + q;
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            const int w = 2 * s + (g >> 1), wi = w < 5 ? w : w - 5;
            const uint32_t vsh = vq[wi * P], vveg = vq[(SW_VIS_WORDS + wi) * P];
            const uint32_t word = !in ? 0u : (w < 5 ? (vsh & vveg) : (vsh & ~vveg));
            const uint32_t half = (word >> (16 * (g & 1))) & 0xffffu;
            if (s % 2 == 0) part[t][s / 2] = half;
            else part[t][s / 2] |= half << 16;
        }
    }
    // the table's fragments: per 8-hour block, step and digit the 64 lanes' 16 bytes in lane
    // order, addressed as the table (uniform) plus a 32-bit byte offset
    const AS1 int8_t *frag = G.diffuse_tab;
    const AS1 double *scl = (const AS1 double *)(G.diffuse_tab + (long long)SW_DIFF_DIGITS * nrows * SW_DIFF_K);
#pragma nounroll
    for (int mt = 0; mt < nrows / 16; ++mt) {
        // opaque per block: the byte expansions are then formed step by step, not hoisted out of
        // the loop for every block (80 more registers)
#pragma unroll
        for (int t = 0; t < 4; ++t) asm volatile("" : "+v"(part[t][0]), "+v"(part[t][1]), "+v"(part[t][2]));
        v4i acc[4][SW_DIFF_DIGITS];
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int d = 0; d < SW_DIFF_DIGITS; ++d) acc[t][d] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            // the step's fragment index made opaque: its loads are issued in step order, not all
            // 20 at the start of the block
            uint32_t fo = (uint32_t)(((mt * 5 + s) * SW_DIFF_DIGITS * 64 + lane) * 16);
            asm volatile("" : "+v"(fo));
            v4i a[SW_DIFF_DIGITS];
#pragma unroll
            for (int d = 0; d < SW_DIFF_DIGITS; ++d) a[d] = *(const AS1 v4i *)(frag + (fo + (uint32_t)(d * 1024)));
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const v4i b = bits_to_bytes((part[t][s / 2] >> (16 * (s % 2))) & 0xffffu);
#pragma unroll
                for (int d = 0; d < SW_DIFF_DIGITS; ++d)
                    acc[t][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[d], b, acc[t][d], 0, 0, 0);
            }
        }
        // rows 16 mt + 4 g + i: hour h = 8 mt + 2 g (i = 0, 1) and h + 1 (i = 2, 3). The integer
        // sum is hi 2^14 + lo with hi = acc3 2^7 + acc2 and lo = acc1 2^7 + acc0, both exact in
        // int32 and in FP32 (below 2^23): one fma rounds it once, and the power-of-two scale is
        // exact (a NaN scale gives NaN, as the reference's 0 x NaN terms do). A row whose values
        // all lie below 2^-99 would have a scale under the FP32 range: its sums become 0.
        const int r0 = 16 * mt + 4 * g, h = 8 * mt + 2 * g;
        float sc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) sc[i] = (float)scl[r0 + i];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int hi = acc[t][3][i] * 128 + acc[t][2][i], lo = acc[t][1][i] * 128 + acc[t][0][i];
                v[i] = fmaf((float)hi, 16384.0f, (float)lo) * sc[i];
            }
            // this block's values are final here: without the barrier all 64 sums are converted
            // at once
            asm volatile("" : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]));
            bool in;
            const long long q = pixel_of(t, in);
            if (!in) continue;
            if (h < n_hours && !(sc[0] < 0.0f)) {
                v2f o;
                o.x = v[0];
                o.y = v[1];
                G.diffuse[(long long)h * P + q] = o;
            }
            if (h + 1 < n_hours && !(sc[2] < 0.0f)) {
                v2f o;
                o.x = v[2];
                o.y = v[3];
                G.diffuse[(long long)(h + 1) * P + q] = o;
            }
        }
    }
}

struct Coef7 {
    double c[7];
};

__global__ void k_utci_scalar(const float *__restrict__ tmrt, long long n, Coef7 cf, float ta,
                              int sentinel, const uint8_t *__restrict__ ground, float *__restrict__ out) {
    const double *c = cf.c;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)blockDim.x * gridDim.x) {
        const float t = tmrt[i];
        if (ground && !ground[i]) { out[i] = __builtin_nanf(""); continue; }
        if (sentinel || t <= -999.0f) { out[i] = -999.0f; continue; }
        const double dT = (double)(t - ta);
        double acc = c[6];
        for (int k = 5; k >= 0; --k) acc = acc * dT + c[k];
        out[i] = (float)acc;
    }
}

// Plain streaming copy: the reference point for HBM byte counters.
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                              long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)blockDim.x * gridDim.x)
        dst[i] = src[i];
}

int launch_hour_shadow(const sw_domain *doms, const int32_t *bs, int nd, int total, int hour,
                       bool full, hipStream_t s, int n_hours = 1, bool per_hour = false) {
    const dim3 grid(total, n_hours);
    if (full)
        hipLaunchKernelGGL(k_hour_shadow<true>, grid, dim3(BLOCK), 0, s, doms, bs, nd, total, hour, (int)per_hour);
    else
        hipLaunchKernelGGL(k_hour_shadow<false>, grid, dim3(BLOCK), 0, s, doms, bs, nd, total, hour, (int)per_hour);
    return hip_check(hipGetLastError(), "k_hour_shadow launch");
}

// Records one workgroup stages for the ground view: the larger of a 16 x 16 and a 32 x 8 tile
// with a halo of second - 1 pixels (wider tiles read the records from global memory), or 0 when
// that exceeds 64 KiB.
int lds_records(int second) {
    const int h = second - 1;
    static_assert(SW_SEARCH_WORDS * 4 <= 65536, "search tables exceed the LDS limit");
    const int a = (16 + 2 * h) * (16 + 2 * h), b = (32 + 2 * h) * (8 + 2 * h);
    const int n = a > b ? a : b;
    return n * (int)sizeof(float) * 2 <= 65536 ? n : 0;
}

// LDS of a pixel-pass workgroup: the staged records during the walk, then the hour's patch search
// tables.
size_t pixel_lds_bytes(int cap) {
    size_t bytes = (size_t)cap * sizeof(float) * 2;
    return bytes < (size_t)SW_SEARCH_WORDS * 4 ? (size_t)SW_SEARCH_WORDS * 4 : bytes;
}

// groups_per_cu > 0: the launch asks for enough LDS per workgroup that at most that many of its
// workgroups share a CU (the LDS is the CU's, so the cap holds whatever the registers allow).
int launch_day_pixels(const sw_domain *doms, const int32_t *bs, int nd, int total, int h0, int h1,
                      const int32_t *gofs, const double *gb, int first, int second, hipStream_t s,
                      int groups_per_cu = 0) {
    const int cap = lds_records(second);
    size_t bytes = pixel_lds_bytes(cap);
    if (groups_per_cu > 0) {
        int dev = 0, per_cu = 0;
        int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
        if (!rc) rc = hip_check(hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev),
                                "hipDeviceGetAttribute");
        if (rc) return rc;
        // one allocation granule above per_cu / (groups_per_cu + 1): groups_per_cu + 1 workgroups
        // no longer fit, groups_per_cu still do
        const size_t need = (size_t)per_cu / (size_t)(groups_per_cu + 1) + 1024;
        if (need * (size_t)groups_per_cu > (size_t)per_cu || need > 65536)
            return fail(SW_ERR_ARGUMENT, "sw_day_pixels_capped: groups_per_cu out of range for this device");
        if (need > bytes) bytes = need;
    }
    hipLaunchKernelGGL(k_day_pixels, dim3(total), dim3(BLOCK), bytes, s, doms, bs, nd, total, h0, h1,
                       gofs, gb, first, second, cap);
    return hip_check(hipGetLastError(), "k_day_pixels launch");
}

int launch_hour_pixels(const sw_domain *doms, const int32_t *bs, int nd, int total, int hour,
                       const int32_t *gofs, const double *gb, int first, int second, bool full,
                       hipStream_t s) {
    const int cap = lds_records(second);
    const size_t bytes = pixel_lds_bytes(cap);
    if (full)
        hipLaunchKernelGGL(k_hour_pixels<true>, dim3(total), dim3(BLOCK), bytes, s, doms, bs, nd, total,
                           hour, gofs, gb, first, second, cap);
    else
        hipLaunchKernelGGL(k_hour_pixels<false>, dim3(total), dim3(BLOCK), bytes, s, doms, bs, nd, total,
                           hour, gofs, gb, first, second, cap);
    return hip_check(hipGetLastError(), "k_hour_pixels launch");
}

int check_launch(const void *doms, const void *bs, int nd, int total) {
    if (!doms || !bs || nd <= 0 || total <= 0) return fail(SW_ERR_ARGUMENT, "empty domain table");
    return SW_OK;
}


// =====================================================================================
// Adjoint of J = (1/N) sum over ground pixels of UTCI at the objective hour, with the shadow
// and sky-view masks held fixed. Per hour (backward from the objective hour):
//   k_adj_prep   per receiver pixel: at the objective hour the forward of the hour is
//                recomputed and the chain UTCI -> Tmrt -> Sstr -> radiation terms is reversed;
//                every hour reverses the lag recursion and the ground-view sums, writing the
//                per-direction coefficients of the walk and adding its own terms;
//   k_adj_patch  (objective hour) per-patch sums of lambda_Kdown / lambda_Kside;
//   k_adj_gather per source pixel: the transposed walk (the 18 x second samples that reach it)
//                and the chain through the neighbour record to the albedo/emissivity maps.
// Per-hour scalar derivatives are summed per wave in FP32 and across waves in FP64.
// =====================================================================================
struct AdjG {
    AS1 float *lam_state, *lam_rec, *grad_alb, *grad_emis, *lam_patch;
    AS1 v4f *coef;
    const AS1 float *svfveg;
    AS1 double *scalars, *patch, *objective, *partial;
};

__device__ __forceinline__ AdjG adjoint_at(const sw_adjoint *a_) {
    const AS4 sw_adjoint *a = (const AS4 sw_adjoint *)a_;
    AdjG g;
    g.lam_state = (AS1 float *)a->lam_state; g.lam_rec = (AS1 float *)a->lam_rec;
    g.coef = (AS1 v4f *)a->coef; g.grad_alb = (AS1 float *)a->grad_alb;
    g.grad_emis = (AS1 float *)a->grad_emis; g.lam_patch = (AS1 float *)a->lam_patch;
    g.svfveg = (const AS1 float *)a->svfveg; g.scalars = (AS1 double *)a->scalars;
    g.patch = (AS1 double *)a->patch; g.objective = (AS1 double *)a->objective;
    g.partial = (AS1 double *)a->partial;
    return g;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums the per-thread values v[0..N) over the workgroup and stores them as the workgroup's
// partial sums of slots slot0 .. slot0 + N in A.partial ([slot][total], one column per
// workgroup); k_adj_reduce adds the columns up. Every thread must call. (Device-scope atomics
// from every wave to the same few addresses serialise: 15.6k waves adding 3 values took 0.56 ms.)
// A record without the work area (partial NULL) gets one atomic add per workgroup and slot into
// the hour's slots or the objective instead.
template <int N>
__device__ __forceinline__ void block_partials(const AdjG &A, int total, int hour, int slot0, const float (&v)[N]) {
    AS1 double *part = A.partial;
    __shared__ float red[BLOCK / 64][N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const float t = wave_sum(v[i]);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][i] = t;
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i < N; i += BLOCK) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) s += (double)red[w][i];
        const int slot = slot0 + i;
        if (part) {
            part[(long long)slot * total + blockIdx.x] = s;
        } else if (s != 0.0) {
            AS1 double *dst = slot < SWA_SLOTS ? A.scalars + (long long)hour * SWA_SLOTS + slot
                                               : A.objective + (slot - SWA_SLOTS);
            atomicAdd((double *)dst, s);
        }
    }
}

__device__ __forceinline__ bool dir_in(int dir, int card) {
    const int az = 5 + 20 * dir;
    switch (card) {
    case 1: return az < 180;                    // east half
    case 2: return az >= 90 && az < 270;        // south half
    case 3: return az >= 180;                   // west half
    default: return az >= 270 || az < 90;       // north half
    }
}

// Reverse of ground_view for one receiver: for each direction the walk is recomputed, its
// coefficients (dJ per unit of a sample's Lup / albedo x shadow / albedo, and the walk length)
// are stored for the gather, and the receiver's own terms (near field at step 0, local terms),
// the wall-flux and wall-albedo derivatives are added. lg, la, ln: adjoints of the ground-view
// sums (all, E, S, W, N) of Lup, albedo x shadow and albedo.
template <bool CHECK>
__device__ __forceinline__ void adj_walk(const DomainG &G, const AdjG &A, const PixelAt &at,
                                         const v4f me, const AS4 double *H, const AS4 int32_t *gofs,
                                         const AS4 double *gbounds, int first, int second,
                                         const float (&lg)[5], const float (&la)[5],
                                         const float (&ln)[5], float &d_lwall, float &d_albedo_b) {
    const int rows = (int)G.rows, cols = (int)G.cols;
    const long long P = G.pixels, p = at.p;
    const float b = (float)(__float_as_uint(me.w) & 0xffu);
    const bool wallbol = G.walls[p] > 0.0f;
    const double aspect = ((double)G.aspect[p] * 3.141592653589793) / 180.0;
    const float firstf = (float)first, secondf = (float)second;
    const int stride = (second + CHUNK - 1) / CHUNK * CHUNK;
    const float k4 = FAR_W, k5 = NEAR_W;
    const float inv_f = 1.0f / firstf, inv_f1 = 1.0f / (firstf + 1.0f);
    const float inv_s = 1.0f / secondf, inv_s1 = 1.0f / (secondf + 1.0f);
    float self_lup = 0.f, self_albsh = 0.f, self_alb = 0.f;
    for (int dir = 0; dir < 18; ++dir) {
        float mu = lg[0], nu = la[0], ka = ln[0];
#pragma unroll
        for (int c = 1; c < 5; ++c)
            if (dir_in(dir, c)) { mu += lg[c]; nu += la[c]; ka += ln[c]; }
        const AS4 int32_t *off = gofs + dir * stride * 2;
        Walk w = {b, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        Walk f = w;
        bool snapped = false;
        long long qprev = p;
        float rep = 0.f;                              // alive steps that repeat qprev off the raster
        for (int s0 = 0; s0 < second; s0 += CHUNK) {
            if (__ballot(w.af != 0.0f) == 0) break;
            v4f buf[CHUNK];
            bool in[CHUNK];
#pragma unroll
            for (int i = 0; i < CHUNK; ++i) {
                const int dr = off[2 * (s0 + i)], dc = off[2 * (s0 + i) + 1];
                long long q;
                if (!CHECK) {
                    in[i] = true;
                    q = p + (long long)(dr * cols + dc);
                } else {
                    const int rr = at.r + dr, cc = at.c + dc;
                    in[i] = (unsigned)rr < (unsigned)rows && (unsigned)cc < (unsigned)cols;
                    q = in[i] ? (long long)rr * cols + cc : qprev;
                    qprev = q;
                }
                buf[i] = record16(G, q);
            }
#pragma unroll
            for (int i = 0; i < CHUNK; ++i) {
                const int st = s0 + i;
                if (st >= second) break;
                const uint32_t fl = __float_as_uint(buf[i].w);
                w.af *= (float)(fl & 0xffu);
                w.L += w.af;
                w.wf = __uint_as_float(max(__float_as_uint(w.wf),
                                           __float_as_uint(w.af * (float)((fl >> 8) & 0xffu))));
                w.cw = fmaf(w.wf, w.af, w.cw);
                if (CHECK && !in[i]) rep += w.af;
                if (st == first - 1) { f = w; snapped = true; }
            }
        }
        if (!snapped) f = w;
        const float cnt_wall = w.cw + w.wf * (secondf - w.L), cnt_any = secondf - w.L;
        const float f_cnt_wall = f.cw + f.wf * (firstf - f.L), f_cnt_any = firstf - f.L;
        const bool sunin2 = cnt_wall > 0.f, f_sunin = f_cnt_wall > 0.f;
        const bool wallin2 = cnt_any > 0.f, f_wallin = f_cnt_any > 0.f;
        const AS4 double *bd = gbounds + dir * 3;
        int facesh;
        if (bd[2] == 0.0)
            facesh = (int)(aspect < bd[0] || aspect >= bd[1]) - (wallbol ? 1 : 0) + 1;
        else
            facesh = 1 - (int)(aspect > bd[0] || aspect <= bd[1]);
        const bool keep = (cnt_wall == secondf) && facesh == 0;
        const float cs = k4 * (sunin2 ? inv_s1 : inv_s);
        const float cf = k5 * (f_sunin ? inv_f1 : inv_f);
        const float a_lup = mu * cs, a_albsh = nu * cs, a_nosh = ka * b * (k4 * inv_s);
        v4f cv;
        cv.x = a_lup; cv.y = a_albsh; cv.z = a_nosh; cv.w = w.L;
        A.coef[(long long)dir * P + p] = cv;
        if (f.L >= 1.0f) {                             // near field: step 0 is the receiver itself
            self_lup += mu * cf;
            self_albsh += nu * cf;
            self_alb += ka * b * (k5 * (f_wallin ? inv_f1 : inv_f));
        }
        self_lup += mu * (1.0f - b);                  // local terms of a non-ground receiver
        self_albsh += nu * (1.0f - b);
        self_alb += ka * (1.0f - b);
        const float wall_s = (sunin2 && !keep) ? cnt_wall * inv_s1 : 0.f;
        const float wall_f = f_sunin ? f_cnt_wall * inv_f1 : 0.f;
        d_lwall += mu * (k4 * wall_s + k5 * wall_f);
        d_albedo_b += nu * (k4 * wall_s + k5 * wall_f) +
                      ka * b * (k4 * (wallin2 ? cnt_any * inv_s : 0.f) +
                                k5 * (f_wallin ? f_cnt_any * inv_f1 : 0.f));
        if (CHECK && rep > 0.f) {                      // repeated off-raster samples land on qprev
            atomicAdd((float *)(A.lam_rec + qprev), a_lup * rep);
            atomicAdd((float *)(A.lam_rec + P + qprev), a_albsh * rep);
            atomicAdd((float *)(A.lam_rec + 2 * P + qprev), a_nosh * rep);
        }
    }
    atomicAdd((float *)(A.lam_rec + p), self_lup);
    atomicAdd((float *)(A.lam_rec + P + p), self_albsh);
    atomicAdd((float *)(A.lam_rec + 2 * P + p), self_alb);
}

// adj_walk for a tile staged in LDS by stage_tile_fixed<TX, GV_FIXED_STEPS - 1, EDGE> (first == 1,
// second == GV_FIXED_STEPS). A direction's walk state is fixed by three numbers: L, the steps
// before the first non-ground sample (static, gvf_len), kw, the first step on a sunlit wall (the
// records' sign bits, read from LDS as ground_view_fixed reads them), and for EDGE tiles E, the
// in-raster steps (static, gvf_edge; later steps repeat step E - 1's sample). adj_walk's running
// state gives, after the walk: sunlit-wall count second - kw when kw < L (else 0), wall count
// second - L, and for the near field (step 0, the receiver) L >= 1 and kw == 0; the receiver's
// facing bit of gvf_len decides ``keep`` as in GvSums::add_far. Only the walked steps that stay
// alive past E (L - E of them) land on the last in-raster sample.
template <int TX, bool EDGE>
__device__ __forceinline__ void adj_walk_fixed(const DomainG &G, const AdjG &A, const PixelAt &at, const bool valid,
                                               const v4f me, const uint32_t me_y, const AS4 int32_t *gofs,
                                               const float (&lg)[5], const float (&la)[5], const float (&ln)[5],
                                               float &d_lwall, float &d_albedo_b) {
    constexpr int NS = GV_FIXED_STEPS, HH = NS - 1, PITCH = TX + 2 * HH;
    constexpr int stride = (NS + CHUNK - 1) / CHUNK * CHUNK;   // padded step axis of gofs
    const long long P = G.pixels, p = at.p;
    const float b = (float)(__float_as_uint(me.w) & 0xffu);
    const int base = (at.r - at.tr0) * PITCH + (at.c - at.tc0);
    uint32_t lw[5], ew[5];
#pragma unroll
    for (int w = 0; w < 5; ++w) {
        lw[w] = G.gvf_len[w * P + p];
        ew[w] = EDGE ? G.gvf_edge[w * P + p] : 0u;
    }
    const float secondf = (float)NS;
    const float k4 = FAR_W, k5 = NEAR_W;
    const float inv_f = 1.0f, inv_f1 = 1.0f / 2.0f;
    const float inv_s = 1.0f / secondf, inv_s1 = 1.0f / (secondf + 1.0f);
    float self_lup = 0.f, self_albsh = 0.f, self_alb = 0.f;
#pragma unroll
    for (int dir = 0; dir < 18; ++dir) {
        float mu = lg[0], nu = la[0], ka = ln[0];
#pragma unroll
        for (int c = 1; c < 5; ++c)
            if (dir_in(dir, c)) { mu += lg[c]; nu += la[c]; ka += ln[c]; }
        const uint32_t lb = (lw[dir >> 2] >> (8 * (dir & 3))) & 0xffu;
        const int L = (int)(lb & 0x7fu);
        const int E = EDGE ? (int)((ew[dir >> 2] >> (8 * (dir & 3))) & 0xffu) : NS;
        uint32_t cm = 0u, last = me_y;
        int nexec = 0;
#pragma unroll
        for (int s0 = 0; s0 < NS; s0 += CHUNK) {
            if (__ballot(s0 < L) == 0) break;
            const int n = NS - s0 < CHUNK ? NS - s0 : CHUNK;
            uint32_t y[CHUNK];
#pragma unroll
            for (int i = 0; i < n; ++i) {
                const int k = (GV_STEPS[dir][s0 + i][0] + HH) * PITCH + GV_STEPS[dir][s0 + i][1] + HH;
                y[i] = __float_as_uint(gv_lds[base + k].y);
            }
#pragma unroll
            for (int i = 0; i < n; ++i) {
                if (EDGE) {
                    const uint32_t keep = s0 + i < E ? 0xffffffffu : 0u;   // bit select, no branch
                    last = (y[i] & keep) | (last & ~keep);
                    y[i] = last;
                }
                cm = __builtin_amdgcn_alignbit(cm, y[i], 31);
            }
            nexec = s0 + n;
        }
        const uint32_t lead = cm == 0u ? 0xffffffffu : (uint32_t)__builtin_clz(cm);
        const uint32_t kw = lead - (uint32_t)(32 - nexec);
        const float Lf = (float)L;
        const bool sunin2 = kw < (uint32_t)L;
        const bool f_alive = L >= 1;                   // the near field's one step (the receiver)
        const bool f_sunin = f_alive && kw == 0u;
        const bool keep = sunin2 && kw == 0u && (lb & 0x80u) != 0u;
        const float cnt_wall = sunin2 ? secondf - (float)kw : 0.f, cnt_any = secondf - Lf;
        const bool wallin2 = cnt_any > 0.f;
        const float f_cnt_any = f_alive ? 0.f : 1.f;
        const float cs = k4 * (sunin2 ? inv_s1 : inv_s);
        const float cf = k5 * (f_sunin ? inv_f1 : inv_f);
        const float a_lup = mu * cs, a_albsh = nu * cs, a_nosh = ka * b * (k4 * inv_s);
        if (valid) {
            v4f cv;
            cv.x = a_lup; cv.y = a_albsh; cv.z = a_nosh; cv.w = Lf;
            A.coef[(long long)dir * P + p] = cv;
        }
        if (f_alive) {                                 // near field: step 0 is the receiver itself
            self_lup += mu * cf;
            self_albsh += nu * cf;
            self_alb += ka * b * (k5 * inv_f);
        }
        self_lup += mu * (1.0f - b);                  // local terms of a non-ground receiver
        self_albsh += nu * (1.0f - b);
        self_alb += ka * (1.0f - b);
        const float wall_s = (sunin2 && !keep) ? cnt_wall * inv_s1 : 0.f;
        const float wall_f = f_sunin ? inv_f1 : 0.f;
        d_lwall += mu * (k4 * wall_s + k5 * wall_f);
        d_albedo_b += nu * (k4 * wall_s + k5 * wall_f) +
                      ka * b * (k4 * (wallin2 ? cnt_any * inv_s : 0.f) + k5 * (f_alive ? 0.f : f_cnt_any * inv_f1));
        if (EDGE && valid && L > E) {                  // alive steps past the raster repeat step E - 1
            const AS4 int32_t *off = gofs + dir * stride * 2;
            const long long q = (long long)(at.r + off[2 * (E - 1)]) * G.cols + (at.c + off[2 * (E - 1) + 1]);
            const float rep = (float)(L - E);
            atomicAdd((float *)(A.lam_rec + q), a_lup * rep);
            atomicAdd((float *)(A.lam_rec + P + q), a_albsh * rep);
            atomicAdd((float *)(A.lam_rec + 2 * P + q), a_nosh * rep);
        }
    }
    if (valid) {
        atomicAdd((float *)(A.lam_rec + p), self_lup);
        atomicAdd((float *)(A.lam_rec + P + p), self_albsh);
        atomicAdd((float *)(A.lam_rec + 2 * P + p), self_alb);
    }
}

// One day hour of k_adj_prep for one pixel. MODE 0-3: the tile (width 16 for 0-1, 32 for 2-3;
// odd: the halo leaves the raster) and its ground-view halo are staged in LDS and the walks read
// them there (first == 1, second == GV_FIXED_STEPS); every thread of the workgroup calls, lanes
// past the raster (valid false) compute a copy of the nearest pixel and write nothing. MODE 4:
// global records, called for valid pixels only.
template <bool TOP, int MODE>
__device__ __forceinline__ void adj_prep_day(const DomainG &G, const AdjG &A, const PixelAt &at, const bool valid,
                                             const AS4 double *H, const AS4 int32_t *gofs,
                                             const AS4 double *gbounds, int first, int second, double inv_n,
                                             float (&sl)[SWA_SLOTS], float (&obj)[2]) {
    constexpr bool STAGED = MODE < 4, EDGE = (MODE & 1) != 0;
    constexpr int TX = MODE < 2 ? 16 : 32;
    if (STAGED) stage_tile_fixed<TX, GV_FIXED_STEPS - 1, EDGE>(G, at);
    const long long P = G.pixels, p = at.p;
    const float emis = G.emis[p];
    const v4f me = record16(G, p);
    const uint32_t cls = (__float_as_uint(me.w) >> 16) & 3u;
    const float shadow = shadow_of_class(cls, (float)(1.0 - (1.0 - H[SWH_PSI])));
    const v2f me2 = G.rec[p];
    const float building = G.ground[p] ? 1.0f : 0.0f;
    const float flux4 = (float)H[SWH_FLUX4];
    const float amb = flux4 * emis;
    const double w = H[SWH_LAG_WEIGHT];
    const bool lf = H[SWH_LAG_FIRST] != 0.0, lu = H[SWH_LAG_UPDATE] != 0.0;
    const int m = second - 1;
    const bool interior = at.r >= m && at.c >= m && at.r < (int)G.rows - m && at.c < (int)G.cols - m;
    float ld[5] = {0.f, 0.f, 0.f, 0.f, 0.f};      // adjoint of the delayed Lup (all, E, S, W, N)
    float la[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, ln[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    if (TOP) {
        // ---- forward of the hour (as k_hour_pixels, day branch) ----------------------
        GvfOut g;
        if (STAGED) g = ground_view_fixed<TX, EDGE>(G, at, me2, building, H);
        else if (interior) g = ground_view<0>(G, at, me2, building, H, gofs, gbounds, first, second);
        else g = ground_view<1>(G, at, me2, building, H, gofs, gbounds, first, second);
        const AS1 float *lg = G.lag;
        float s0 = lg[p], s1 = lg[P + p], s2 = lg[2 * P + p], s3 = lg[3 * P + p], s4 = lg[4 * P + p];
        const float Lup = lag(g.all.lup * INV18 + amb, s0, w, lf, lu);
        const float LupE = lag(g.e.lup * INV9 + amb, s1, w, lf, lu);
        const float LupS = lag(g.s.lup * INV9 + amb, s2, w, lf, lu);
        const float LupW = lag(g.w.lup * INV9 + amb, s3, w, lf, lu);
        const float LupN = lag(g.n.lup * INV9 + amb, s4, w, lf, lu);
        const float radI = (float)H[SWH_RADI], radD = (float)H[SWH_RADD], radG = (float)H[SWH_RADG];
        const float albedo_b = (float)H[SWH_ALBEDO_B];
        const float Fsh = wedge(G.tan_svfalfa[p], H[SWH_TANZEN]);
        const float svfbuveg = G.svfbuveg[p];
        const float sinalt = (float)H[SWH_SINALT], cosalt = (float)H[SWH_COSALT];
        const float X = radG * (1.0f - Fsh) + radD * Fsh;
        const float around = albedo_b * (1.0f - svfbuveg) * X;
        const float refl = radD * svfbuveg + around;
        const float Kup = (g.all.alb * INV18) * radI * sinalt + refl * (g.all.nosh * INV18);
        const float KupE = (g.e.alb * INV9) * radI * sinalt + refl * (g.e.nosh * INV9);
        const float KupS = (g.s.alb * INV9) * radI * sinalt + refl * (g.s.nosh * INV9);
        const float KupW = (g.w.alb * INV9) * radI * sinalt + refl * (g.w.nosh * INV9);
        const float KupN = (g.n.alb * INV9) * radI * sinalt + refl * (g.n.nosh * INV9);
        const float KsideI = shadow * radI * cosalt;
        const int row = uniform_index(H[SWH_PATCH_TABLE]);
        const AS4 float *pf = G.ptab_f + (long long)row * SW_PATCHES * SW_PATCH_F;
        const AS4 double *pd = G.ptab_d + (long long)row * SW_PATCH_D * SW_PATCHES;
        float a_lv = 0.f, a_kd = 0.f, a_swsun = 0.f, a_swsh = 0.f, a_lsunc = 0.f, a_lsuns = 0.f;
        float a_lshc = 0.f, a_lshs = 0.f;
        day_patch_sums<true>(G.vis, P, p, pf, pd, H, G.tan_asvf[p], (float)H[SWH_DIFF_VEG], a_lv, a_kd,
                       a_swsun, a_swsh, a_lsunc, a_lsuns, a_lshc, a_lshs);
        const AS1 float *ss = G.skysums;
        const float SA = (float)H[SWH_SKY_A], SB = (float)H[SWH_SKY_B];
        const float ss0 = ss[SWS_SKY_SIN * P + p], ss1 = ss[SWS_SKY_GSIN * P + p];
        const float ss2 = ss[SWS_SKY_COS * P + p], ss3 = ss[SWS_SKY_GCOS * P + p];
        const float a_ldn = SA * ss0 - SB * ss1, a_lsd = SA * ss2 - SB * ss3;
        const float a_vc = ss[SWS_VEG_COS * P + p], a_vs = ss[SWS_VEG_SIN * P + p];
        const float a_rc = ss[SWS_OBS_COS * P + p], a_rs = ss[SWS_OBS_SIN * P + p];
        const float dRad = a_lv * radD;
        const float Kdown = radI * shadow * sinalt + around + dRad;
        const float sw_sun = (float)H[SWH_SW_SUNLIT], sw_sh = (float)H[SWH_SW_SHADED];
        const float Kside = KsideI + a_kd + sw_sun * a_swsun + sw_sh * a_swsh + sw_sh * a_vc;
        const float lw_amb = (float)H[SWH_LW_AMBIENT], lw_sun = (float)H[SWH_LW_SUNLIT];
        const float lw_refl = (float)H[SWH_LW_REFL];
        const float reflsurf = (a_ldn + Lup) * lw_refl;
        const float Ldown = a_ldn + lw_amb * a_vs + lw_amb * a_lshs + lw_sun * a_lsuns + reflsurf * a_rs;
        const float Lside = a_lsd + lw_amb * a_vc + lw_amb * a_lshc + lw_sun * a_lsunc + reflsurf * a_rc;
        const float absK = (float)H[SWH_ABSK], absL = (float)H[SWH_ABSL];
        const float Fup = (float)H[SWH_FUP], Fside = (float)H[SWH_FSIDE], Fcyl = (float)H[SWH_FCYL];
        const float Ksum = (KupN + KupE + KupS + KupW) * 0.5f;
        const float Lsum = (LupN + LupE + LupS + LupW) * 0.5f;
        const float Sstr = absK * (Kside * Fcyl + (Kdown + Kup) * Fup + Ksum * Fside) +
                           absL * ((Ldown + Lup) * Fup + Lside * Fcyl + Lsum * Fside);
        const float Tmrt = tmrt_of(Sstr, absL);
        const bool ground = G.ground[p] != 0;
        // ---- reverse ---------------------------------------------------------------------
        float lU = 0.f, lT = 0.f;
        if (ground && H[SWH_UTCI0 + 6] != -999.0 && Tmrt > -999.0f) {
            const double dT = (double)(Tmrt - (float)H[SWH_TA32]);
            double acc = H[SWH_UTCI0 + 6], der = 6.0 * H[SWH_UTCI0 + 6];
            for (int k = 5; k >= 0; --k) {
                acc = acc * dT + H[SWH_UTCI0 + k];
                if (k >= 1) der = der * dT + (double)k * H[SWH_UTCI0 + k];
            }
            obj[0] = (float)acc;
            obj[1] = 1.0f;
            lU = (float)inv_n;
            lT = (float)(inv_n * der);
            double pw = 1.0;
            for (int k = 0; k < 7; ++k) { sl[SWA_UTCI0 + k] = (float)(inv_n * pw); pw *= dT; }
            sl[SWA_TA32] = -lT;
        }
        (void)lU;
        const float lS = lT * (Tmrt + 273.2f) / (4.0f * Sstr);
        const float lKside = lS * absK * Fcyl, lKdown = lS * absK * Fup, lKup = lS * absK * Fup;
        const float lKupC = lS * absK * Fside * 0.5f;
        const float lLdown = lS * absL * Fup, lLside = lS * absL * Fcyl;
        const float lLupC = lS * absL * Fside * 0.5f;
        const float lrefls = lLdown * a_rs + lLside * a_rc;
        ld[0] = lS * absL * Fup + lrefls * lw_refl;
        ld[1] = ld[2] = ld[3] = ld[4] = lLupC;
        const float l_ldn = lLdown + lrefls * lw_refl;
        sl[SWA_SKY_A] = l_ldn * ss0 + lLside * ss2;
        sl[SWA_SKY_B] = -(l_ldn * ss1 + lLside * ss3);
        sl[SWA_LW_AMBIENT] = lLdown * (a_vs + a_lshs) + lLside * (a_vc + a_lshc);
        sl[SWA_LW_SUNLIT] = lLdown * a_lsuns + lLside * a_lsunc;
        sl[SWA_SW_SUNLIT] = lKside * a_swsun;
        sl[SWA_SW_SHADED] = lKside * (a_swsh + a_vc);
        if (valid) {
            A.lam_patch[p] = lKdown;
            A.lam_patch[P + p] = lKside;
        }
        // Kup and the ground-view albedo sums
        const float gsum_alb = g.all.alb * INV18, gsum_nosh = g.all.nosh * INV18;
        float lrefl = lKup * gsum_nosh;
        float d_radI = lKup * sinalt * gsum_alb;
        la[0] = lKup * radI * sinalt * INV18;
        ln[0] = lKup * refl * INV18;
        const Triple gc[4] = {g.e, g.s, g.w, g.n};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            la[c + 1] = lKupC * radI * sinalt * INV9;
            ln[c + 1] = lKupC * refl * INV9;
            lrefl += lKupC * (gc[c].nosh * INV9);
            d_radI += lKupC * sinalt * (gc[c].alb * INV9);
        }
        const float laround = lKdown + lrefl;
        d_radI += lKdown * shadow * sinalt + lKside * shadow * cosalt;
        sl[SWA_RADI] = d_radI;
        sl[SWA_RADD] = lKdown * a_lv + lrefl * svfbuveg + laround * albedo_b * (1.0f - svfbuveg) * Fsh;
        sl[SWA_RADG] = laround * albedo_b * (1.0f - svfbuveg) * (1.0f - Fsh);
        sl[SWA_ALBEDO_B] = laround * (1.0f - svfbuveg) * X;
        const float lsvfbuveg = lrefl * radD - laround * albedo_b * X;
        sl[SWA_TRANS_STATIC] = lsvfbuveg * (1.0f - A.svfveg[p]);
        const float lshadow = lKdown * radI * sinalt + lKside * radI * cosalt;
        sl[SWA_PSI] = cls == 1u ? lshadow : 0.f;
    }
    // ---- lag recursion, reversed (the delayed value is used this hour, the state by the next)
    AS1 float *ls = A.lam_state;
    float lcur[5];
    const float wf = (float)w;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const float lso = ls[k * P + p];
        const float ldt = ld[k] + (lu ? lso : 0.f);
        const float lprev = ldt * wf + (lu ? 0.f : lso);
        lcur[k] = ldt * (1.0f - wf) + (lf ? lprev : 0.f);
        if (valid) ls[k * P + p] = lf ? 0.f : lprev;
    }
    // current Lup = ground-view sum / 18 (or / 9 per half) + flux4 * emis
    const float lg5[5] = {lcur[0] * INV18, lcur[1] * INV9, lcur[2] * INV9, lcur[3] * INV9, lcur[4] * INV9};
    const float lamb = lcur[0] + lcur[1] + lcur[2] + lcur[3] + lcur[4];
    if (valid) A.grad_emis[p] += lamb * flux4;
    sl[SWA_FLUX4] = lamb * emis;
    float d_lwall = 0.f, d_albedo_b = 0.f;
    if (STAGED)
        adj_walk_fixed<TX, EDGE>(G, A, at, valid, me, __float_as_uint(me2.y), gofs, lg5, la, ln, d_lwall,
                                 d_albedo_b);
    else if (interior)
        adj_walk<false>(G, A, at, me, H, gofs, gbounds, first, second, lg5, la, ln, d_lwall, d_albedo_b);
    else
        adj_walk<true>(G, A, at, me, H, gofs, gbounds, first, second, lg5, la, ln, d_lwall, d_albedo_b);
    sl[SWA_LWALL] = d_lwall;
    sl[SWA_ALBEDO_B] += d_albedo_b;
}

template <bool TOP>
__global__ __launch_bounds__(BLOCK) void k_adj_prep(const sw_domain *__restrict__ doms,
                                                    const int32_t *__restrict__ bs, int total, int hour,
                                                    const sw_adjoint *__restrict__ adj_,
                                                    const int32_t *__restrict__ gofs_,
                                                    const double *__restrict__ gbounds_, int first,
                                                    int second, double inv_n, int lds_cap) {
    const AS4 int32_t *gofs = (const AS4 int32_t *)gofs_;
    const AS4 double *gbounds = (const AS4 double *)gbounds_;
    DomainG G;
    PixelAt at = locate(doms, bs, 1, total, G);
    const bool valid = at.valid;
    if (!valid) {
        at.r = min(at.r, (int)G.rows - 1);
        at.c = min(at.c, (int)G.cols - 1);
        at.p = (long long)at.r * G.cols + at.c;
    }
    const AdjG A = adjoint_at(adj_);
    const AS4 double *H = G.hours + (long long)hour * SW_HOUR_STRIDE;
    float sl[SWA_SLOTS];
#pragma unroll
    for (int i = 0; i < SWA_SLOTS; ++i) sl[i] = 0.f;
    float obj[2] = {0.f, 0.f};
    if (H[SWH_DAY] != 0.0) {
        const int m = second - 1;
        const bool fits = lds_cap > 0 && at.tx <= 32 && (at.tx + 2 * m) * (at.ty + 2 * m) <= lds_cap;
        const bool inner = at.tr0 >= m && at.tc0 >= m && at.tr0 + at.ty + m <= (int)G.rows &&
                           at.tc0 + at.tx + m <= (int)G.cols;
        const bool fixed = first == 1 && second == GV_FIXED_STEPS && (at.tx == 16 || at.tx == 32) &&
                           (inner || G.gvf_edge != nullptr);
        if (fits && fixed) {
            if (at.tx == 16 && inner)
                adj_prep_day<TOP, 0>(G, A, at, valid, H, gofs, gbounds, first, second, inv_n, sl, obj);
            else if (at.tx == 16)
                adj_prep_day<TOP, 1>(G, A, at, valid, H, gofs, gbounds, first, second, inv_n, sl, obj);
            else if (inner)
                adj_prep_day<TOP, 2>(G, A, at, valid, H, gofs, gbounds, first, second, inv_n, sl, obj);
            else
                adj_prep_day<TOP, 3>(G, A, at, valid, H, gofs, gbounds, first, second, inv_n, sl, obj);
        } else if (valid) {
            adj_prep_day<TOP, 4>(G, A, at, true, H, gofs, gbounds, first, second, inv_n, sl, obj);
        }
        if (!valid) {
#pragma unroll
            for (int i = 0; i < SWA_SLOTS; ++i) sl[i] = 0.f;
            obj[0] = obj[1] = 0.f;
        }
    }
    float all[SWA_PARTIAL_SLOTS];
#pragma unroll
    for (int i = 0; i < SWA_SLOTS; ++i) all[i] = sl[i];
    all[SWA_SLOTS] = obj[0];
    all[SWA_SLOTS + 1] = obj[1];
    block_partials(A, total, hour, 0, all);
}

// Per-patch sums of lambda_Kdown and lambda_Kside over the pixels that see the patch as sky or
// through vegetation only (objective hour).
// One visibility word w (blockIdx.y) per launch row: each thread sums its pixels' lambda_Kdown and
// lambda_Kside over the word's 32 patches in registers (grid-stride over the pixels), then each
// sum is reduced over the workgroup and added once per workgroup. The pixel count and buffers are
// read from the device records, so the launch needs nothing from the host.
constexpr int ADJ_PATCH_BLOCKS = 128;
__global__ __launch_bounds__(256) void k_adj_patch(const sw_domain *__restrict__ dom,
                                                   const sw_adjoint *__restrict__ adj_) {
    const DomainG G = domain_at(dom);
    const AdjG A = adjoint_at(adj_);
    const AS1 uint32_t *bits = G.vis;
    const AS1 float *lam = A.lam_patch;
    AS1 double *out = A.patch;
    const long long P = G.pixels;
    const int w = (int)blockIdx.y, k0 = w * 32;
    float sK[32], vK[32], sS[32], vS[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) sK[j] = vK[j] = sS[j] = vS[j] = 0.f;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x) {
        const float lK = lam[p], lS = lam[P + p];
        const uint32_t vsh = bits[(0 * SW_VIS_WORDS + w) * P + p];
        const uint32_t vveg = bits[(1 * SW_VIS_WORDS + w) * P + p];
        const uint32_t sky = vsh & vveg, vegonly = vsh & ~vveg;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const bool a = (sky >> j) & 1u, v = (vegonly >> j) & 1u;
            sK[j] += a ? lK : 0.f;
            vK[j] += v ? lK : 0.f;
            sS[j] += a ? lS : 0.f;
            vS[j] += v ? lS : 0.f;
        }
    }
    const int n = SW_PATCHES - k0 < 32 ? SW_PATCHES - k0 : 32;
    __shared__ float red[4][4][32];                   // [wave][sum][patch of the word]
    const int wave = (int)threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        if (j >= n) break;
        const float t0 = wave_sum(sK[j]), t1 = wave_sum(vK[j]), t2 = wave_sum(sS[j]), t3 = wave_sum(vS[j]);
        if ((threadIdx.x & 63) == 0) {
            red[wave][0][j] = t0;
            red[wave][1][j] = t1;
            red[wave][2][j] = t2;
            red[wave][3][j] = t3;
        }
    }
    __syncthreads();
    if (threadIdx.x < 128) {
        const int kind = (int)threadIdx.x >> 5, j = (int)threadIdx.x & 31;
        static_assert(SWA_PATCH_SKY_KDOWN == 0 && SWA_PATCH_VEG_KDOWN == 1 && SWA_PATCH_SKY_KSIDE == 2 &&
                      SWA_PATCH_VEG_KSIDE == 3, "patch sum order");
        if (j < n) {
            const float t = (red[0][kind][j] + red[1][kind][j]) + (red[2][kind][j] + red[3][kind][j]);
            if (t != 0.f) atomicAdd((double *)(out + kind * SW_PATCHES + k0 + j), (double)t);
        }
    }
}

// Receivers whose walk in direction DIR reaches a TX-wide tile of sources: the sample of step s
// of receiver p lands on q = p + shift(DIR, s), so the receivers of the tile's sources form the
// tile grown by the step shifts' extent on the far side (a box of nr x nc pixels whose corner
// is (r0, c0) from the tile's corner).
struct GatherBox {
    int r0, c0, nr, nc;
};
constexpr GatherBox gather_box(int dir, int tx) {
    int rmin = 0, rmax = 0, cmin = 0, cmax = 0;
    for (int s = 0; s < GV_FIXED_STEPS; ++s) {
        const int dr = GV_STEPS[dir][s][0], dc = GV_STEPS[dir][s][1];
        rmin = dr < rmin ? dr : rmin; rmax = dr > rmax ? dr : rmax;
        cmin = dc < cmin ? dc : cmin; cmax = dc > cmax ? dc : cmax;
    }
    return GatherBox{-rmax, -cmax, BLOCK / tx + rmax - rmin, tx + cmax - cmin};
}
constexpr int gather_box_max(int tx) {
    int n = 0;
    for (int d = 0; d < 18; ++d) {
        const GatherBox b = gather_box(d, tx);
        n = b.nr * b.nc > n ? b.nr * b.nc : n;
    }
    return n;
}
// LDS of k_adj_gather's staged form: two coefficient boxes (one direction is read while the next
// is staged), for the larger of the 16- and 32-wide tiles.
constexpr int GATHER_BOX_RECORDS = gather_box_max(16) > gather_box_max(32) ? gather_box_max(16) : gather_box_max(32);
constexpr int GATHER_LDS_BYTES = 2 * GATHER_BOX_RECORDS * (int)sizeof(v4f);
static_assert(GATHER_LDS_BYTES <= 65536, "gather boxes exceed the LDS limit");

// One direction of the transposed walk for a tile staged box by box: the receivers' coefficients
// of direction DIR (zero for receivers outside the raster) are copied into LDS buffer DIR & 1,
// then each source sums the coefficients of the receivers whose step s lands on it while alive
// (s < L), in the order of the global form (steps ascending).
template <int TX, int DIR>
__device__ __forceinline__ void gather_dir(const DomainG &G, const AdjG &A, const PixelAt &at, float &l_lup,
                                          float &l_albsh, float &l_alb) {
    constexpr GatherBox bx = gather_box(DIR, TX);
    constexpr int N = bx.nr * bx.nc, IT = (N + BLOCK - 1) / BLOCK;
    v4f *buf = reinterpret_cast<v4f *>(gv_lds) + (DIR & 1) * GATHER_BOX_RECORDS;
    const int rows = (int)G.rows, cols = (int)G.cols;
    const AS1 v4f *cd = A.coef + (long long)DIR * G.pixels;
    // opaque per direction: otherwise the box positions of all 18 directions are computed up
    // front and held (about 180 VGPRs)
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    v4f v[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int idx = min(t + BLOCK * it, N - 1);
        const int gr = at.tr0 + bx.r0 + idx / bx.nc, gc = at.tc0 + bx.c0 + idx % bx.nc;
        const bool in = (unsigned)gr < (unsigned)rows && (unsigned)gc < (unsigned)cols;
        v[it] = cd[in ? (long long)gr * cols + gc : 0];
        if (!in) v[it] = v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) buf[min(t + BLOCK * it, N - 1)] = v[it];
    __syncthreads();
    const int base = (at.r - at.tr0 - bx.r0) * bx.nc + (at.c - at.tc0 - bx.c0);
#pragma unroll
    for (int s = 0; s < GV_FIXED_STEPS; ++s) {
        const v4f x = buf[base - GV_STEPS[DIR][s][0] * bx.nc - GV_STEPS[DIR][s][1]];
        const float on = (float)s < x.w ? 1.0f : 0.0f;
        l_lup = fmaf(on, x.x, l_lup);
        l_albsh = fmaf(on, x.y, l_albsh);
        l_alb = fmaf(on, x.z, l_alb);
        // at most CHUNK reads in flight, and no direction's loads moved into another's
        if (s % CHUNK == CHUNK - 1 || s == GV_FIXED_STEPS - 1) __builtin_amdgcn_sched_barrier(0);
    }
}

template <int TX, int DIR = 0>
__device__ __forceinline__ void gather_fixed(const DomainG &G, const AdjG &A, const PixelAt &at, float &l_lup,
                                            float &l_albsh, float &l_alb) {
    gather_dir<TX, DIR>(G, A, at, l_lup, l_albsh, l_alb);
    if constexpr (DIR + 1 < 18) gather_fixed<TX, DIR + 1>(G, A, at, l_lup, l_albsh, l_alb);
}

// Transposed walk for each source pixel q: the sample of receiver p = q - shift(d, s) at step s
// reaches q when p's walk is alive there (s < L_d(p)). Then the neighbour record of q is
// reversed to the albedo and emissivity maps and the hour's Ta, CI_TgG and canopy shadow.
// Tiles 16 or 32 wide at second == GV_FIXED_STEPS stage the receivers' coefficients in LDS
// (gather_fixed; every thread of the workgroup takes part, lanes past the raster compute a copy
// of the nearest pixel and write nothing); other tiles gather them from global memory.
__global__ __launch_bounds__(BLOCK) void k_adj_gather(const sw_domain *__restrict__ doms,
                                                      const int32_t *__restrict__ bs, int total,
                                                      int hour, const sw_adjoint *__restrict__ adj_,
                                                      const int32_t *__restrict__ gofs_, int second,
                                                      int lds_bytes) {
#pragma clang fp contract(off)
    const AS4 int32_t *gofs = (const AS4 int32_t *)gofs_;
    DomainG G;
    PixelAt at = locate(doms, bs, 1, total, G);
    const bool valid = at.valid;
    if (!valid) {
        at.r = min(at.r, (int)G.rows - 1);
        at.c = min(at.c, (int)G.cols - 1);
        at.p = (long long)at.r * G.cols + at.c;
    }
    const AdjG A = adjoint_at(adj_);
    const AS4 double *H = G.hours + (long long)hour * SW_HOUR_STRIDE;
    static_assert(SWA_TA == 0 && SWA_PSI == 1 && SWA_CITGG == 2, "gather slots");
    float sl[3] = {0.f, 0.f, 0.f};                    // Ta, psi, CI_TgG
    if (H[SWH_DAY] != 0.0) {
        const int rows = (int)G.rows, cols = (int)G.cols;
        const long long P = G.pixels, q = at.p;
        AS1 float *lr = A.lam_rec;
        float l_lup = lr[q], l_albsh = lr[P + q], l_alb = lr[2 * P + q];
        const bool staged = second == GV_FIXED_STEPS && (at.tx == 16 || at.tx == 32) &&
                            lds_bytes >= GATHER_LDS_BYTES;
        if (staged) {
            if (at.tx == 16) gather_fixed<16>(G, A, at, l_lup, l_albsh, l_alb);
            else gather_fixed<32>(G, A, at, l_lup, l_albsh, l_alb);
        } else if (valid) {
            const int stride = (second + CHUNK - 1) / CHUNK * CHUNK;
            const int m = second - 1;
            const bool interior = at.r >= m && at.c >= m && at.r < rows - m && at.c < cols - m;
            for (int dir = 0; dir < 18; ++dir) {
                const AS4 int32_t *off = gofs + dir * stride * 2;
                const AS1 v4f *cd = A.coef + (long long)dir * P;
                for (int s0 = 0; s0 < second; s0 += CHUNK) {
                    v4f buf[CHUNK];
                    bool in[CHUNK];
#pragma unroll
                    for (int i = 0; i < CHUNK; ++i) {
                        const int dr = off[2 * (s0 + i)], dc = off[2 * (s0 + i) + 1];
                        long long pr;
                        if (interior) {
                            in[i] = true;
                            pr = q - (long long)(dr * cols + dc);
                        } else {
                            const int rr = at.r - dr, cc = at.c - dc;
                            in[i] = (unsigned)rr < (unsigned)rows && (unsigned)cc < (unsigned)cols;
                            pr = in[i] ? (long long)rr * cols + cc : q;
                        }
                        buf[i] = cd[pr];
                    }
#pragma unroll
                    for (int i = 0; i < CHUNK; ++i) {
                        const int st = s0 + i;
                        if (st >= second) break;
                        const float on = (in[i] && (float)st < buf[i].w) ? 1.0f : 0.0f;
                        l_lup = fmaf(on, buf[i].x, l_lup);
                        l_albsh = fmaf(on, buf[i].y, l_albsh);
                        l_alb = fmaf(on, buf[i].z, l_alb);
                    }
                }
            }
        }
        if (valid) {
            lr[q] = 0.f; lr[P + q] = 0.f; lr[2 * P + q] = 0.f;
            // the record of q: Lup = SBC emis ((Tg sh + Ta + 273.15)^4 - (Ta + 273.15)^4), alb, alb sh
            const uint32_t cls = class_of(G.masks[q]);
            const float sh = shadow_of_class(cls, (float)(1.0 - (1.0 - H[SWH_PSI])));
            const double amp = (double)G.tgk[q] * H[SWH_ALTMAX] + (double)G.tstart[q];
            const double base = amp * tg_sine(H, G.tmaxlst[q]);
            double tg = base * H[SWH_CITGG];
            const bool clamped = G.landcover == 1 && tg < 0.0;
            if (clamped) tg = 0.0;
            const double ta = H[SWH_TA], emis = (double)G.emis[q];
            const double tk = tg * (double)sh + ta + 273.15, ta4 = ta + 273.15;
            const double ll = (double)l_lup;
            A.grad_emis[q] += (float)(ll * SBC * ((tk * tk) * (tk * tk) - (ta4 * ta4) * (ta4 * ta4)));
            A.grad_alb[q] += l_alb + l_albsh * sh;
            const double dtk = ll * SBC * emis * 4.0 * (tk * tk * tk);
            sl[0] = (float)(dtk - ll * SBC * emis * 4.0 * (ta4 * ta4 * ta4));
            sl[2] = clamped ? 0.f : (float)(dtk * (double)sh * base);
            sl[1] = cls == 1u ? (float)(dtk * tg) + l_albsh * G.alb[q] : 0.f;
        }
    }
    block_partials(A, total, hour, SWA_TA, sl);
}

// Adds the workgroup partial sums of the slots in mask (one workgroup per slot) to the hour's
// record (slots below SWA_SLOTS) or to the objective sum and count.
__global__ __launch_bounds__(256) void k_adj_reduce(const sw_adjoint *__restrict__ adj_, int total, int hour,
                                                    uint32_t mask) {
    const int slot = (int)blockIdx.x;
    if (!((mask >> slot) & 1u)) return;
    const AdjG A = adjoint_at(adj_);
    if (!A.partial) return;                           // the workgroups added their sums themselves
    const AS1 double *src = A.partial + (long long)slot * total;
    double s = 0.0;
    for (int i = (int)threadIdx.x; i < total; i += 256) s += src[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double t = (red[0] + red[1]) + (red[2] + red[3]);
        AS1 double *dst = slot < SWA_SLOTS ? A.scalars + (long long)hour * SWA_SLOTS + slot
                                           : A.objective + (slot - SWA_SLOTS);
        *dst += t;
    }
}

// =====================================================================================
// Per-step sub-models as standalone elementwise kernels (model.py's cylindric_wedge,
// TsWaveDelay_2015a, Kup_veg_2015a, Lside_veg_v2022a, surface_temperature, assemble_tmrt).
// Each input is a plane of n values or a single value broadcast to all pixels; the field table
// travels by value in the kernel arguments. R is the working precision (float or double).
// =====================================================================================
constexpr int SUB_FIELDS = 20;
struct FieldSet {
    const void *ptr[SUB_FIELDS];
    long long count[SUB_FIELDS];
};

template <typename T>
__device__ __forceinline__ T fld(const FieldSet &F, int k, long long i) {
    return ((const T *)F.ptr[k])[F.count[k] == 1 ? 0 : i];
}

// Sunlit wall fraction of a cylinder, umep_solweig.py:383-426, in FP64 without the NaN
// replacement of the step (the caller of the sub-model sees NaN where the reference does).
__device__ __forceinline__ double wedge64(double alfa, double zen) {
    const double PI = 3.141592653589793;
    const double ta = tan(alfa), tb = tan(zen);
    const double xa = 1.0 - 2.0 / (ta * tb), ha = 2.0 / (ta * tb), ba = 1.0 / ta;
    double ukil = 0.0;
    if (xa < 0.0) {
        const double qa = tb / 2.0;
        const double za = sqrt(ba * ba - (qa * qa) / 4.0);
        const double phi = atan(za / qa);
        ukil = 2.0 * ba * xa * ((sin(phi) - phi * cos(phi)) / (1.0 - cos(phi)));
    }
    return (2.0 * PI * ba - (2.0 * ba * ha + ukil)) / (2.0 * PI * ba);
}

template <typename R>
__global__ void k_sub_wedge(FieldSet F, long long n, double zen, R *out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (R)wedge64(fld<double>(F, 0, i), zen);
}

template <typename R>
__global__ void k_sub_delay(FieldSet F, long long n, int first, int update, R weight, R *out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const R cur = fld<R>(F, 0, i);
    const R prev = first ? cur : fld<R>(F, 1, i);
    const R d = cur * (R(1) - weight) + prev * weight;
    out[i] = d;
    out[n + i] = update ? d : prev;
}

// fields: svfbuveg, F_sh, gvfalb (all, E, S, W, N), gvfalbnosh (all, E, S, W, N)
template <typename R>
__global__ void k_sub_kup(FieldSet F, long long n, R radI, R radD, R radG, double altitude, R albedo_b, R *out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const R svfb = fld<R>(F, 0, i), fsh = fld<R>(F, 1, i);
    const R s = (R)sin(altitude * (3.141592653589793 / 180.0));
    const R refl = radD * svfb + albedo_b * (R(1) - svfb) * (radG * (R(1) - fsh) + radD * fsh);
    for (int c = 0; c < 5; ++c) out[c * n + i] = fld<R>(F, 2 + c, i) * radI * s + refl * fld<R>(F, 7 + c, i);
}

template <typename R>
__device__ __forceinline__ R view_poly(R x) {
    return R(63.227) * pow(x, R(6)) - R(161.51) * pow(x, R(5)) + R(156.91) * pow(x, R(4)) -
           R(70.424) * pow(x, R(3)) + R(16.773) * pow(x, R(2)) - R(0.4863) * x;
}

// fields (faces E, S, W, N): svf[4], svfveg[4], svfaveg[4], Tw, Ldown, F_sh, Lup[4]
template <typename R>
__global__ void k_sub_lside(FieldSet F, long long n, double azimuth, double altitude, R Ta, R SBC, R ewall,
                            R esky, double t, R CI, int anisotropic, R *out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const R total = R(4.4897);
    const R T4 = pow(Ta + R(273.15), R(4));
    const R c = R(1) - CI;
    const R allsky = esky * SBC * T4 * (R(1) - c) + c * SBC * T4;
    const R Fs = R(2) * fld<R>(F, 14, i) - R(1);
    const R Tw = fld<R>(F, 12, i), Ldown = fld<R>(F, 13, i);
    for (int f = 0; f < 4; ++f) {
        const R up = fld<R>(F, 15 + f, i);
        if (anisotropic == 1) { out[f * n + i] = up * R(0.5); continue; }
        const R svf = fld<R>(F, f, i), veg = fld<R>(F, 4 + f, i), aveg = fld<R>(F, 8 + f, i);
        const R wwall = (total - view_poly(svf)) / total - (total - view_poly(aveg)) / total;
        const R both = veg + svf - R(1);
        const R wsky = view_poly(both) / total, wrefl = (total - view_poly(both)) / total;
        const R wveg = wrefl - wwall;
        // face f lit by the sun, and the sun's azimuth relative to the face
        bool lit;
        double rel;
        switch (f) {
        case 0: lit = azimuth > 180.0 - t && azimuth <= 360.0 - t; rel = azimuth - 180.0 + t; break;
        case 1: lit = azimuth <= 90.0 - t || azimuth > 270.0 - t; rel = azimuth - 270.0 + t; break;
        case 2: lit = azimuth > 360.0 - t || azimuth <= 180.0 - t; rel = azimuth + t; break;
        default: lit = azimuth > 90.0 - t && azimuth <= 270.0 - t; rel = azimuth - 90.0 + t; break;
        }
        R sun = R(0), shade = SBC * ewall * T4 * wwall * R(0.5);
        if (altitude > 0.0 && lit) {
            const R alfa = asin(exp(log(R(1) - svf) / R(2)));
            const R ab = atan(alfa), bb = atan(tan(alfa * Fs));
            const R betasun = (ab - bb) / R(2) + bb;
            const R tw = Ta + R(273.15) + Tw * (R)sin(rel * (3.141592653589793 / 180.0));
            sun = SBC * ewall * pow(tw, R(4)) * wwall * (R(1) - Fs) * cos(betasun) * R(0.5);
            shade = SBC * ewall * T4 * wwall * Fs * R(0.5);
        }
        const R sky = ((svf + veg - R(1)) * allsky) * wsky * R(0.5);
        const R vegl = SBC * ewall * T4 * wveg * R(0.5);
        const R refl = (Ldown + up) * wrefl * (R(1) - ewall) * R(0.5);
        out[f * n + i] = sky + sun + shade + vegl + up * R(0.5) + refl;
    }
}

// fields: TgK, Tstart, TgK_wall, Tstart_wall, TmaxLST, TmaxLST_wall
template <typename R>
__global__ void k_sub_surface(FieldSet F, long long n, double altmax, double frac, double snup, R citgg,
                              int landcover, R *out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const R PI = R(3.141592653589793);
    const R x0 = R(frac) - R(snup) / R(24);
    R tg = (fld<R>(F, 0, i) * R(altmax) + fld<R>(F, 1, i)) *
           sin((x0 / (fld<R>(F, 4, i) / R(24) - R(snup) / R(24))) * PI / R(2));
    R tw = (fld<R>(F, 2, i) * R(altmax) + fld<R>(F, 3, i)) *
           sin((x0 / (fld<R>(F, 5, i) / R(24) - R(snup) / R(24))) * PI / R(2));
    if (tw < R(0)) tw = R(0);
    tg = tg * citgg;
    tw = tw * citgg;
    if (landcover == 1 && tg < R(0)) tg = R(0);
    out[i] = tg;
    out[n + i] = tw;
}

// fields: Kdown, Kup, Knorth, Keast, Ksouth, Kwest, Ldown, Lup, Lnorth, Least, Lsouth, Lwest,
//         KsideI, Kside, Lside
template <typename R>
__global__ void k_sub_assemble(FieldSet F, long long n, R absK, R absL, R Fside, R Fup, R Fcyl, int cyl,
                               int anisotropic, R SBC, R *out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const R K4 = fld<R>(F, 2, i) + fld<R>(F, 3, i) + fld<R>(F, 4, i) + fld<R>(F, 5, i);
    const R L4 = fld<R>(F, 8, i) + fld<R>(F, 9, i) + fld<R>(F, 10, i) + fld<R>(F, 11, i);
    const R KdKu = fld<R>(F, 0, i) + fld<R>(F, 1, i), LdLu = fld<R>(F, 6, i) + fld<R>(F, 7, i);
    R S;
    if (cyl == 1 && anisotropic == 1)
        S = absK * (fld<R>(F, 13, i) * Fcyl + KdKu * Fup + K4 * Fside) +
            absL * (LdLu * Fup + fld<R>(F, 14, i) * Fcyl + L4 * Fside);
    else if (cyl == 1)
        S = absK * (fld<R>(F, 12, i) * Fcyl + KdKu * Fup + K4 * Fside) + absL * (LdLu * Fup + L4 * Fside);
    else
        S = absK * (KdKu * Fup + K4 * Fside) + absL * (LdLu * Fup + L4 * Fside);
    out[i] = sqrt(sqrt(S / (absL * SBC))) - R(273.2);
    out[n + i] = S;
}

// Host side: one launcher per sub-model, dispatching on the working precision.
// The elementwise terms of one step that the reference evaluates as CuPy expressions
// (model.py:120-126, 165-171, 183-195), in the working precision R with contraction off and the
// reference's operation order:
//   op 0  shadow = sh - (1 - vegsh)(1 - psi)        fields sh, vegsh (FP64)     params psi
//   op 1  Kdown = radI shadow sin(alt) + dRad + albedo_b (1 - svfbuveg)(radG (1 - F_sh) + radD F_sh)
//                                                   fields shadow, dRad, svfbuveg, F_sh
//                                                   params radI, sin(alt), albedo_b, radG, radD
//   op 2  TgTemp = Tg shadow + Ta                   fields Tg, shadow           params Ta
//   op 3  night Lup = SBC emis (Ta + 273.15)^4, SBC 0.98 (Twater + 273.15)^4 on water with land cover
//                                                   fields emis, lc_grid (FP64) params Ta, SBC, landcover, water Lup
//   op 4  x, with NaN replaced by v                 fields x                    params v
//   op 5  a + b (the cardinal longwave sums, model.py:244-251)  fields a, b     no params
//   op 6  x rounded to R (utci_calculator's float32 Tmrt, utci.py:8-13)  fields x (FP64)  no params
constexpr int TERMS_OPS = 7;
constexpr int TERMS_NFIELDS[TERMS_OPS] = {2, 4, 2, 2, 1, 2, 1};
constexpr int TERMS_NPARAMS[TERMS_OPS] = {1, 5, 1, 4, 1, 0, 0};

template <typename R>
__global__ void k_step_terms(int op, FieldSet F, long long n, Coef7 q, R *out) {
#pragma clang fp contract(off)
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    R v;
    switch (op) {
    case 0: v = (R)(fld<double>(F, 0, i) - (1.0 - fld<double>(F, 1, i)) * (1.0 - q.c[0])); break;
    case 1: {
        const R shadow = fld<R>(F, 0, i), dRad = fld<R>(F, 1, i), svfbuveg = fld<R>(F, 2, i), fsh = fld<R>(F, 3, i);
        const R direct = (R)q.c[0] * shadow * (R)q.c[1] + dRad;
        v = direct + (R)q.c[2] * (R(1) - svfbuveg) * ((R)q.c[3] * (R(1) - fsh) + (R)q.c[4] * fsh);
        break;
    }
    case 2: v = fld<R>(F, 0, i) * fld<R>(F, 1, i) + (R)q.c[0]; break;
    case 3: {
        const R t = (R)q.c[0] + R(273.15);
        v = (R)q.c[1] * fld<R>(F, 0, i) * pow(t, R(4));
        if ((int)q.c[2] == 1 && fld<double>(F, 1, i) == 3.0) v = (R)q.c[3];
        break;
    }
    case 4: {
        const R x = fld<R>(F, 0, i);
        v = isnan(x) ? (R)q.c[0] : x;
        break;
    }
    case 5: v = fld<R>(F, 0, i) + fld<R>(F, 1, i); break;
    default: v = (R)fld<double>(F, 0, i); break;
    }
    out[i] = v;
}

// dRad = radD x sum_k diffsh[k][p] lum[k] (the reference's diffuse kernel, shortwave.cu:98-107): the
// products in FP64, the running sum and the result rounded to R.
template <typename R>
__global__ void k_diffuse(const R *__restrict__ vis, const double *__restrict__ lum, long long n, int patches,
                          double radD, R *__restrict__ out) {
#pragma clang fp contract(off)
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    R sum = R(0);
    for (int k = 0; k < patches; ++k) sum = sum + (R)((double)vis[(long long)k * n + p] * lum[k]);
    out[p] = (R)((double)sum * radD);
}

template <typename R>
int launch_sub(int which, const FieldSet &F, long long n, const double *q, void *out, hipStream_t s) {
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    R *o = (R *)out;
    switch (which) {
    case 0: hipLaunchKernelGGL(k_sub_wedge<R>, grid, block, 0, s, F, n, q[0], o); break;
    case 1: hipLaunchKernelGGL(k_sub_delay<R>, grid, block, 0, s, F, n, (int)q[0], (int)q[1], (R)q[2], o); break;
    case 2: hipLaunchKernelGGL(k_sub_kup<R>, grid, block, 0, s, F, n, (R)q[0], (R)q[1], (R)q[2], q[3], (R)q[4], o);
        break;
    case 3: hipLaunchKernelGGL(k_sub_lside<R>, grid, block, 0, s, F, n, q[0], q[1], (R)q[2], (R)q[3], (R)q[4],
                               (R)q[5], q[6], (R)q[7], (int)q[8], o); break;
    case 4: hipLaunchKernelGGL(k_sub_surface<R>, grid, block, 0, s, F, n, q[0], q[1], q[2], (R)q[3], (int)q[4], o);
        break;
    default: hipLaunchKernelGGL(k_sub_assemble<R>, grid, block, 0, s, F, n, (R)q[0], (R)q[1], (R)q[2], (R)q[3],
                                (R)q[4], (int)q[5], (int)q[6], (R)q[7], o); break;
    }
    return hip_check(hipGetLastError(), "sub-model launch");
}

// fields expected and parameters expected per sub-model
constexpr int SUB_NFIELDS[6] = {1, 2, 12, 19, 6, 15};
constexpr int SUB_NPARAMS[6] = {1, 3, 5, 9, 5, 8};

int make_fieldset(const void *const *fields, const long long *counts, int n_fields, long long n, FieldSet &F) {
    for (int k = 0; k < n_fields; ++k) {
        if (!fields[k] || (counts[k] != 1 && counts[k] != n))
            return fail(SW_ERR_ARGUMENT, "sub-model: a field is NULL or neither 1 nor n values");
        F.ptr[k] = fields[k];
        F.count[k] = counts[k];
    }
    return SW_OK;
}

int run_sub(int which, const void *const *fields, const long long *counts, int n_fields, long long n,
            const double *params, int n_params, int real_bytes, void *out, void *stream) {
    if (!fields || !counts || !params || !out) return fail(SW_ERR_ARGUMENT, "sub-model: NULL pointer");
    if (n_fields != SUB_NFIELDS[which] || n_params != SUB_NPARAMS[which])
        return fail(SW_ERR_ARGUMENT, "sub-model: wrong number of fields or parameters");
    if (real_bytes != 4 && real_bytes != 8) return fail(SW_ERR_ARGUMENT, "sub-model: precision is 4 or 8 bytes");
    if (n <= 0) return SW_OK;
    FieldSet F = {};
    const int rc = make_fieldset(fields, counts, n_fields, n, F);
    if (rc) return rc;
    return real_bytes == 8 ? launch_sub<double>(which, F, n, params, out, (hipStream_t)stream)
                           : launch_sub<float>(which, F, n, params, out, (hipStream_t)stream);
}

// =====================================================================================
// Standalone ground view factors for arbitrary caller fields (model.gvf, model.py:569-622;
// gvf_2018a / sunonsurface_2018a, umep_solweig.py:583-663, 1400-1600). Unlike the fused hour
// pass, the ground mask may hold any value (the walk keeps its running minimum) and the shadow
// may be fractional, so the walk follows the reference step by step in the working precision
// R, with contraction off so the FP64 form rounds as NumPy does.
// Fields: 0 buildings, 1 walls, 2 wall aspect (radians), 3 wallsun, 9 land cover (all FP64);
//         4 shadow, 5 Tg, 6 Tgwall, 7 emis, 8 albedo (R).
// =====================================================================================
template <typename R>
struct GvfSample {
    R bu, sh, sunwall, lup0, lup1, albsh, alb, pad;
};

// Per pixel: the neighbour sample of every walk and the Tg after the water replacement
// (umep_solweig.py:1419-1421). lup0 uses the caller's Tg (the first direction, which runs before
// the in-place replacement), lup1 the replaced one (the other 17 directions).
template <typename R>
__global__ void k_gvf_pack(FieldSet F, long long n, R ta, R t273, R sbc, double t4a, int landcover,
                           double twater_minus_ta, GvfSample<R> *rec, R *tg_out) {
#pragma clang fp contract(off)
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double bu = fld<double>(F, 0, i);
    const R sh = fld<R>(F, 4, i), tg = fld<R>(F, 5, i), emis = fld<R>(F, 7, i), alb = fld<R>(F, 8, i);
    R tg1 = tg;
    if (landcover == 1 && fld<double>(F, 9, i) == 3.0) tg1 = (R)twater_minus_ta;
    const R se = sbc * emis, amb = se * (R)t4a;
    GvfSample<R> s;
    s.bu = (R)bu;
    s.sh = sh;
    s.sunwall = (fld<double>(F, 3, i) / fld<double>(F, 1, i)) * bu == 1.0 ? R(1) : R(0);
    s.lup0 = se * pow(tg * sh + ta + t273, R(4)) - amb;
    s.lup1 = se * pow(tg1 * sh + ta + t273, R(4)) - amb;
    s.albsh = alb * sh;
    s.alb = alb;
    s.pad = R(0);
    rec[i] = s;
    tg_out[i] = tg1;
}

// Sums of one direction's walk; ``sn`` is the snapshot after the first ``first`` steps.
template <typename R>
struct GvfSums {
    R wall, sh, lwall, lupsh, albwall, albsh, albwallnosh, albnosh;
};

template <typename R>
__device__ __forceinline__ R gvf_mix(R wall_sum, R ground_sum, R wall_div, R ground_div, R has_wall) {
#pragma clang fp contract(off)
    return ((wall_sum + ground_sum) / wall_div) * has_wall + ground_sum / ground_div * (has_wall * R(-1) + R(1));
}

// One thread per receiver pixel: 18 directions x ``second`` steps, a step whose shift leaves the
// raster repeating the previous sample (the shifted-window assignment of the reference).
// Output planes (17 + updated Tg written by the pack kernel): gvfLup, gvfalb, gvfalbnosh, then
// the E, S, W, N triples, gvfSum, gvfNorm.
template <typename R>
__global__ void __launch_bounds__(256) k_gvf_walk(FieldSet F, const GvfSample<R> *rec, int rows, int cols,
                                                  const int32_t *offsets, int stride, const double *bounds,
                                                  int first, int second, R ta, R t273, R sbc_ewall,
                                                  R sbc_ewall_t4a, R albedo_b, R ambient_sbc, R t4a, R *out) {
#pragma clang fp contract(off)
    const long long n = (long long)rows * cols;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int r = (int)(p / cols), c = (int)(p % cols);
    const GvfSample<R> me = rec[p];
    const bool wallbol = fld<double>(F, 1, p) > 0.0;
    const double aspect = fld<double>(F, 2, p);
    const R lwall = sbc_ewall * pow(fld<R>(F, 6, p) + ta + t273, R(4)) - sbc_ewall_t4a;
    const R emis = fld<R>(F, 7, p);
    const R firstf = (R)first, secondf = (R)second;
    R acc_lup = 0, acc_alb = 0, acc_nosh = 0, acc_sum = 0;
    R card[4][3] = {};
    for (int dir = 0; dir < 18; ++dir) {
        const int32_t *off = offsets + (long long)dir * stride * 2;
        const bool d0 = dir == 0;
        GvfSample<R> t = me;
        R f = me.bu;
        GvfSums<R> w = {}, s = {};
        R bub = 0, bubwall = 0;
        for (int st = 0; st < second; ++st) {
            const int rr = r + off[2 * st], cc = c + off[2 * st + 1];
            if ((unsigned)rr < (unsigned)rows && (unsigned)cc < (unsigned)cols) t = rec[(long long)rr * cols + cc];
            if (t.bu < f) f = t.bu;
            w.sh = w.sh + t.sh * f;
            w.lupsh = w.lupsh + (d0 ? t.lup0 : t.lup1) * f;
            w.albsh = w.albsh + t.albsh * f;
            w.albnosh = w.albnosh + t.alb * f;
            const R tempb = t.sunwall * f;
            const R tempbwall = f * R(-1) + R(1);
            bub = (tempb + bub) > R(0) ? R(1) : R(0);
            bubwall = (tempbwall + bubwall) > R(0) ? R(1) : R(0);
            w.lwall = w.lwall + bub * lwall;
            w.albwall = w.albwall + bub * albedo_b;
            w.wall = w.wall + bub;
            w.albwallnosh = w.albwallnosh + bubwall * albedo_b;
            if (st < first) s = w;
        }
        const R sunin = s.wall > R(0) ? R(1) : R(0), wallin = s.albwallnosh > R(0) ? R(1) : R(0);
        const R sunin2 = w.wall > R(0) ? R(1) : R(0), wallin2 = w.albwallnosh > R(0) ? R(1) : R(0);
        const double *bd = bounds + dir * 3;
        int facesh;
        if (bd[2] == 0.0)
            facesh = (int)(aspect < bd[0] || aspect >= bd[1]) - (wallbol ? 1 : 0) + 1;
        else
            facesh = 1 - (int)(aspect > bd[0] || aspect <= bd[1]);
        const bool keep = (w.wall == secondf) && facesh == 0;
        const R f1 = firstf + R(1), s1 = secondf + R(1);
        if (keep) { w.wall = 0; w.lwall = 0; w.albwall = 0; }
        R gvf2 = gvf_mix(w.wall, w.sh, s1, secondf, sunin2);
        if (gvf2 > R(1)) gvf2 = R(1);
        const R lup1 = gvf_mix(s.lwall, s.lupsh, f1, firstf, sunin);
        const R lup2 = gvf_mix(w.lwall, w.lupsh, s1, secondf, sunin2);
        const R alb1 = gvf_mix(s.albwall, s.albsh, f1, firstf, sunin);
        const R alb2 = gvf_mix(w.albwall, w.albsh, s1, secondf, sunin2);
        const R nosh1 = gvf_mix(s.albwallnosh, s.albnosh, f1, firstf, wallin);
        const R nosh2 = gvf_mix(w.albwallnosh, w.albnosh, secondf, secondf, wallin2);
        const R notbu = me.bu * R(-1) + R(1);
        const R lup = (lup1 * R(0.5) + lup2 * R(0.4)) / R(0.9) + me.lup1 * notbu;
        const R alb = (alb1 * R(0.5) + alb2 * R(0.4)) / R(0.9) + me.alb * notbu * me.sh;
        const R nosh = (nosh1 * R(0.5) + nosh2 * R(0.4)) / R(0.9) * me.bu + me.alb * notbu;
        acc_lup = acc_lup + lup;
        acc_alb = acc_alb + alb;
        acc_nosh = acc_nosh + nosh;
        acc_sum = acc_sum + gvf2;
        const int az = 5 + 20 * dir;
        const bool in_card[4] = {az < 180, az >= 90 && az < 270, az >= 180, az >= 270 || az < 90};
        for (int k = 0; k < 4; ++k)
            if (in_card[k]) {
                card[k][0] = card[k][0] + lup;
                card[k][1] = card[k][1] + alb;
                card[k][2] = card[k][2] + nosh;
            }
    }
    const R ambient = ambient_sbc * emis * t4a;
    out[p] = acc_lup / R(18) + ambient;
    out[n + p] = acc_alb / R(18);
    out[2 * n + p] = acc_nosh / R(18);
    for (int k = 0; k < 4; ++k) {
        out[(3 + 3 * k) * n + p] = card[k][0] / R(9) + ambient;
        out[(4 + 3 * k) * n + p] = card[k][1] / R(9);
        out[(5 + 3 * k) * n + p] = card[k][2] / R(9);
    }
    out[15 * n + p] = acc_sum;
    out[16 * n + p] = me.bu == R(0) ? R(1) : acc_sum / R(18);
}

// params: Ta, ewall, SBC, albedo_b, landcover, Twater
template <typename R>
int launch_gvf(const FieldSet &F, int rows, int cols, const int32_t *offsets, int stride, const double *bounds,
               int first, int second, const double *q, void *work, void *out, hipStream_t s) {
    const long long n = (long long)rows * cols;
    const double ta = q[0], ewall = q[1], sbc = q[2], albedo_b = q[3], twater = q[5];
    const int landcover = (int)q[4];
    const double t4a = pow(ta + 273.15, 4.0);
    GvfSample<R> *rec = (GvfSample<R> *)work;
    R *o = (R *)out;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_gvf_pack<R>, grid, block, 0, s, F, n, (R)ta, (R)273.15, (R)sbc, t4a, landcover,
                       twater - ta, rec, o + 17 * n);
    hipLaunchKernelGGL(k_gvf_walk<R>, grid, block, 0, s, F, rec, rows, cols, offsets, stride, bounds, first,
                       second, (R)ta, (R)273.15, (R)(sbc * ewall), (R)(sbc * ewall * t4a), (R)albedo_b, (R)sbc,
                       (R)t4a, o);
    return hip_check(hipGetLastError(), "gvf launch");
}

// =====================================================================================
// Anisotropic-sky sub-models for caller masks (model.prepare_geometry, longwave,
// prepare_shortwave and Kside_veg_v2022a, model.py:273-506). The three [rows, cols, K] masks of a
// patch geometry are packed once into four flag bit planes [4][W][P], W = ceil(K / 32): bit
// k % 32 of word [f][k / 32][p] is flag f of patch k at pixel p, with f = 0 sky (sh == 1 and
// veg == 1), 1 vegetation (veg == 0 or vb == 0), 2 wall ((1 - sh) vb == 1), 3 obstructed (sh == 0
// or veg == 0 or vb == 0): the comparisons of umep_solweig.py:666-911 and 1205-1359. Per-patch
// coefficients are host FP64 tables built for each call; the pixel kernels sum over the patches
// in UMEP's order with contraction off, so the FP64 form rounds as the NumPy code does.
// =====================================================================================
template <typename M>
__global__ void k_pack_flags(const M *sh, const M *veg, const M *vb, long long n, int patches, int words,
                             uint32_t *flags) {
#pragma clang fp contract(off)
    using A = typename std::conditional<std::is_integral<M>::value, int, M>::type;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * words) return;
    const int w = (int)(i / n);
    const long long p = i - (long long)w * n;
    uint32_t fs = 0u, fv = 0u, fw = 0u, fo = 0u;
    const int k0 = w * 32, k1 = min(patches, k0 + 32);
    for (int k = k0; k < k1; ++k) {
        const long long q = p * patches + k;
        const A s = (A)sh[q], v = (A)veg[q], b = (A)vb[q];
        const uint32_t bit = 1u << (k - k0);
        if (s == A(1) && v == A(1)) fs |= bit;
        if (v == A(0) || b == A(0)) fv |= bit;
        if ((A(1) - s) * b == A(1)) fw |= bit;
        if (s == A(0) || v == A(0) || b == A(0)) fo |= bit;
    }
    flags[(0LL * words + w) * n + p] = fs;
    flags[(1LL * words + w) * n + p] = fv;
    flags[(2LL * words + w) * n + p] = fw;
    flags[(3LL * words + w) * n + p] = fo;
}

// Longwave patch table columns (one FP64 row per patch): sky Ldown / Lside of the patch;
// vegetation and shaded-wall terms (surface x steradian x cos or sin of the patch altitude);
// steradian, cos, sin; the cardinal direction cosines E, W, N, S (0 outside the face's sector);
// yi of the wall classification, its two thresholds, and the eligibility of the sun test.
enum { LW_LDN, LW_LSD, LW_VEG_C, LW_VEG_S, LW_SH_C, LW_SH_S, LW_STER, LW_COS, LW_SIN, LW_DE, LW_DW, LW_DN, LW_DS,
       LW_YI, LW_THR_A, LW_THR_B, LW_ELIG, LW_COLS = 20 };

// fields: Lup, Tgwall (R), tan(asvf) (FP64). Output planes Ldown, Lside, Least, Lwest, Lnorth,
// Lsouth (Lcyl_v2022a + define_patch_characteristics, umep_solweig.py:946-1029, 1205-1359).
template <typename R>
__global__ void __launch_bounds__(256) k_longwave(const uint32_t *flags, int words, long long n, int patches,
                                                  const double *tab, FieldSet F, R ta, R t273, R ewall_sbc,
                                                  R one_minus_ewall, R *out) {
#pragma clang fp contract(off)
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const R PI = R(3.141592653589793);
    const R lup = fld<R>(F, 0, p);
    const double tan_asvf = fld<double>(F, 2, p);
    const R ss = (ewall_sbc * pow(ta + fld<R>(F, 1, p) + t273, R(4))) / PI;
    R d_sky = 0, s_sky = 0, d_veg = 0, s_veg = 0, d_sun = 0, s_sun = 0, d_sh = 0, s_sh = 0;
    R E = 0, W = 0, N = 0, S = 0;
    for (int w = 0; w < words; ++w) {
      const uint32_t f_sky = flags[(0LL * words + w) * n + p], f_veg = flags[(1LL * words + w) * n + p];
      const uint32_t f_wall = flags[(2LL * words + w) * n + p];
      const int k1 = min(patches, 32 * w + 32);
      for (int k = 32 * w; k < k1; ++k) {
        const int b = k - 32 * w;
        const double *c = tab + (long long)k * LW_COLS;
        const R dE = (R)c[LW_DE], dW = (R)c[LW_DW], dN = (R)c[LW_DN], dS = (R)c[LW_DS];
        if ((f_sky >> b) & 1u) {
            const R lsd = (R)c[LW_LSD];
            d_sky = d_sky + (R)c[LW_LDN];
            s_sky = s_sky + lsd;
            E = E + lsd * dE; W = W + lsd * dW; N = N + lsd * dN; S = S + lsd * dS;
        }
        if ((f_veg >> b) & 1u) {
            const R vc = (R)c[LW_VEG_C];
            s_veg = s_veg + vc;
            d_veg = d_veg + (R)c[LW_VEG_S];
            E = E + vc * dE; W = W + vc * dW; N = N + vc * dN; S = S + vc * dS;
        }
        if ((f_wall >> b) & 1u) {
            bool sun = false, shd = true;
            if (c[LW_ELIG] != 0.0) {
                const double x = tan_asvf + c[LW_YI];
                sun = x < c[LW_THR_A];
                shd = x >= c[LW_THR_B];
            }
            if (sun) {
                const R st = ss * (R)c[LW_STER];
                const R sc = st * (R)c[LW_COS];
                s_sun = s_sun + sc;
                d_sun = d_sun + st * (R)c[LW_SIN];
                E = E + sc * dE; W = W + sc * dW; N = N + sc * dN; S = S + sc * dS;
            }
            if (shd) {
                const R hc = (R)c[LW_SH_C];
                s_sh = s_sh + hc;
                d_sh = d_sh + (R)c[LW_SH_S];
                E = E + hc * dE; W = W + hc * dW; N = N + hc * dN; S = S + hc * dS;
            }
        }
      }
    }
    const R refl = (((d_sky + lup) * one_minus_ewall) * R(0.5)) / PI;
    R d_ref = 0, s_ref = 0;
    for (int k = 0; k < patches; ++k) {
        if (!((flags[(3LL * words + (k >> 5)) * n + p] >> (k & 31)) & 1u)) continue;
        const double *c = tab + (long long)k * LW_COLS;
        const R st = refl * (R)c[LW_STER];
        const R sc = st * (R)c[LW_COS];
        s_ref = s_ref + sc;
        d_ref = d_ref + st * (R)c[LW_SIN];
        E = E + sc * (R)c[LW_DE]; W = W + sc * (R)c[LW_DW]; N = N + sc * (R)c[LW_DN]; S = S + sc * (R)c[LW_DS];
    }
    out[p] = d_sky + d_veg + d_sh + d_sun + d_ref;
    out[n + p] = s_sky + s_veg + s_sh + s_sun + s_ref;
    out[2 * n + p] = E;
    out[3 * n + p] = W;
    out[4 * n + p] = N;
    out[5 * n + p] = S;
}

// Shortwave patch table columns: lumChi, the cylinder incidence cos(alt) cos(0), steradian, the
// box incidences E, S, W, N; sunlit- and shaded-surface terms (x steradian x cos alt); the
// reflection direction cosines E, S, W, N; yi and the two thresholds; sector flags (bits 0-3:
// diffuse sectors E, S, W, N; 4-7 reflection sectors; 8: the box's sun test applies).
enum { KS_LUMCHI, KS_ANGC, KS_STER, KS_ANG, KS_SUN_C = KS_ANG + 4, KS_SH_C, KS_DIRC, KS_YI = KS_DIRC + 4, KS_THR_A,
       KS_THR_B, KS_FLAGS, KS_COLS = 20 };

// fields: shadow, F_sh, KupE, KupS, KupW, KupN, svfE, svfS, svfW, svfN, svfEveg, svfSveg, svfWveg,
// svfNveg (R), tan(asvf) (FP64). Output planes Keast, Ksouth, Kwest, Knorth, KsideI, KsideD,
// Kside (Kside_veg_v2022a, umep_solweig.py:666-911).
struct KsideParams {
    double radI, radD, radG, cos_alt, one_minus_psi, albedo, sin_face[4];
    int lit[4], cyl, ani;
};

template <typename R>
__global__ void __launch_bounds__(256) k_kside(const uint32_t *flags, int words, const R *diffsh, long long n,
                                               int patches, const double *tab, FieldSet F, KsideParams q,
                                               R *out) {
#pragma clang fp contract(off)
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const R radI = (R)q.radI, radD = (R)q.radD, radG = (R)q.radG, cos_alt = (R)q.cos_alt;
    const R albedo = (R)q.albedo;
    const R shadow = fld<R>(F, 0, p), fsh = fld<R>(F, 1, p);
    R direct[4] = {0, 0, 0, 0}, KsideI;
    if (q.cyl == 1) {
        KsideI = shadow * radI * cos_alt;
    } else {
        for (int f = 0; f < 4; ++f)
            if (q.lit[f]) direct[f] = radI * shadow * cos_alt * (R)q.sin_face[f];
        KsideI = shadow * R(0);
    }
    R K[4], KsideD = 0, Kside = 0;
    if (q.ani != 1) {
        const R total = R(4.4897);
        for (int f = 0; f < 4; ++f) {
            const R svf = fld<R>(F, 6 + f, p), veg = fld<R>(F, 10 + f, p);
            const R wall = (total - view_poly(svf)) / total;
            const R vegw = (total - view_poly(veg + svf - R(1))) / total - wall;
            const R w = wall + vegw * (R)q.one_minus_psi;
            K[f] = direct[f] + (radD * (R(1) - w) + albedo * (w * (radG * (R(1) - fsh) + radD * fsh)) +
                                fld<R>(F, 2 + f, p)) * R(0.5);
        }
    } else {
        const double tan_asvf = fld<double>(F, 14, p);
        R dRad[4] = {0, 0, 0, 0}, rveg[4] = {0, 0, 0, 0}, rsun[4] = {0, 0, 0, 0}, rsh[4] = {0, 0, 0, 0};
        R kveg = 0, ksun = 0, ksh = 0;
        for (int k = 0; k < patches; ++k) {
            const double *c = tab + (long long)k * KS_COLS;
            const int fl = (int)c[KS_FLAGS];
            const R ds = diffsh[(long long)k * n + p];
            if (q.cyl == 1) {
                KsideD = KsideD + ds * (R)c[KS_LUMCHI] * (R)c[KS_ANGC] * (R)c[KS_STER];
            } else {
                for (int f = 0; f < 4; ++f)
                    if ((fl >> f) & 1) dRad[f] = dRad[f] + ds * (R)c[KS_LUMCHI] * (R)c[KS_ANG + f] * (R)c[KS_STER];
            }
            const R shc = (R)c[KS_SH_C], sunc = (R)c[KS_SUN_C];
            const uint32_t bit = 1u << (k & 31);
            if (flags[(1LL * words + (k >> 5)) * n + p] & bit) {
                kveg = kveg + shc;
                if (q.cyl != 1)
                    for (int f = 0; f < 4; ++f)
                        if ((fl >> (4 + f)) & 1) rveg[f] = rveg[f] + shc * (R)c[KS_DIRC + f];
            }
            if (!(flags[(2LL * words + (k >> 5)) * n + p] & bit)) continue;
            const bool test = q.cyl == 1 || ((fl >> 8) & 1);
            bool sun = false, shd = true;
            if (test) {
                const double x = tan_asvf + c[KS_YI];
                sun = x < c[KS_THR_A];
                shd = x >= c[KS_THR_B];
            }
            if (q.cyl == 1) {
                if (sun) ksun = ksun + sunc;
                if (shd) ksh = ksh + shc;
                continue;
            }
            for (int f = 0; f < 4; ++f) {
                if (!((fl >> (4 + f)) & 1)) continue;
                const R d = (R)c[KS_DIRC + f];
                if (sun) rsun[f] = rsun[f] + sunc * d;
                if (shd) rsh[f] = rsh[f] + shc * d;
            }
        }
        for (int f = 0; f < 4; ++f) {
            const R half = fld<R>(F, 2 + f, p) * R(0.5);
            K[f] = q.cyl == 1 ? half : direct[f] + dRad[f] + rsun[f] + rsh[f] + rveg[f] + half;
        }
        if (q.cyl == 1) Kside = KsideI + KsideD + ksun + ksh + kveg;
    }
    for (int f = 0; f < 4; ++f) out[f * n + p] = K[f];
    out[4 * n + p] = KsideI;
    out[5 * n + p] = KsideD;
    out[6 * n + p] = Kside;
}

int launch_pack(const void *sh, const void *veg, const void *vb, int mask_bytes, long long n, int patches,
                uint32_t *flags, hipStream_t s) {
    const int words = (patches + 31) / 32;
    const long long total = n * words;
    const dim3 grid((unsigned)((total + 255) / 256)), block(256);
    switch (mask_bytes) {
    case 1: hipLaunchKernelGGL(k_pack_flags<uint8_t>, grid, block, 0, s, (const uint8_t *)sh, (const uint8_t *)veg,
                               (const uint8_t *)vb, n, patches, words, flags); break;
    case 4: hipLaunchKernelGGL(k_pack_flags<float>, grid, block, 0, s, (const float *)sh, (const float *)veg,
                               (const float *)vb, n, patches, words, flags); break;
    default: hipLaunchKernelGGL(k_pack_flags<double>, grid, block, 0, s, (const double *)sh, (const double *)veg,
                                (const double *)vb, n, patches, words, flags); break;
    }
    return hip_check(hipGetLastError(), "k_pack_flags launch");
}

// params: Ta, ewall, SBC
template <typename R>
int launch_longwave(const uint32_t *flags, int patches, long long n, const double *tab, const FieldSet &F,
                    const double *q, void *out, hipStream_t s) {
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_longwave<R>, grid, block, 0, s, flags, (patches + 31) / 32, n, patches, tab, F, (R)q[0],
                       (R)273.15, (R)(q[1] * q[2]), (R)(1.0 - q[1]), (R *)out);
    return hip_check(hipGetLastError(), "k_longwave launch");
}

template <typename R>
int launch_kside(const uint32_t *flags, const void *diffsh, int patches, long long n, const double *tab,
                 const FieldSet &F, const KsideParams &q, void *out, hipStream_t s) {
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_kside<R>, grid, block, 0, s, flags, (patches + 31) / 32, (const R *)diffsh, n, patches,
                       tab, F, q, (R *)out);
    return hip_check(hipGetLastError(), "k_kside launch");
}

// =====================================================================================
// Wall heights and aspect from a DSM (calculate_dsm_walls, walls.py:24-142; geometry.cu:15-63).
// Heights: the largest edge neighbour minus the pixel where that is >= threshold; on the first
// row the reference's column test (geometry.cu:30-33) keeps -dsm for columns rows..cols-2.
// Aspect: a wall pixel away from the border takes the first of 180 rotated line templates that
// covers the most wall pixels, and faces the lower of the two halves of the rotated two-sided
// template; pixels left at 0 take the DSM gradient direction (np.gradient, walls.py:125-131).
// Templates are host tables of (row, col) offsets.
// =====================================================================================
__global__ void k_wall_heights(const double *dsm, int rows, int cols, double threshold, double *walls,
                               uint8_t *occupancy) {
    const long long n = (long long)rows * cols;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int r = (int)(p / cols), c = (int)(p % cols);
    double h = 0.0;
    if (r > 0 && r < rows - 1 && c > 0 && c < cols - 1) {
        const double nb = fmax(fmax(dsm[p - cols], dsm[p + cols]), fmax(dsm[p - 1], dsm[p + 1]));
        const double d = nb - dsm[p];
        if (d >= threshold) h = d;
    } else if (r == 0 && rows > 1 && c >= rows && c < cols - 1) {
        const double d = -dsm[p];
        if (d >= threshold) h = d;
    }
    walls[p] = h;
    occupancy[p] = h > 0.0 ? 1 : 0;
}

__global__ void k_wall_aspect(const double *dsm, const double *walls, const uint8_t *occupancy, int rows,
                              int cols, int radius, int n_angles, const int32_t *line, const int32_t *line_start,
                              const int32_t *side, const int32_t *side_start, double step, double *aspect) {
#pragma clang fp contract(off)
    const long long n = (long long)rows * cols;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (!(walls[p] > 0.0)) { aspect[p] = 0.0; return; }
    const int r = (int)(p / cols), c = (int)(p % cols);
    double a = 0.0;
    if (n_angles > 0 && r >= radius && r < rows - radius - 2 && c >= radius && c < cols - radius - 2) {
        int best = 0, angle = -1;
        for (int k = 0; k < n_angles; ++k) {
            int score = 0;
            for (int j = line_start[k]; j < line_start[k + 1]; ++j)
                score += occupancy[p + (long long)line[2 * j] * cols + line[2 * j + 1]];
            if (score > best) { best = score; angle = k; }
        }
        if (angle >= 0) {
            const int32_t *ss = side_start + 3 * angle;
            double first = 0.0, second = 0.0;
            for (int j = ss[0]; j < ss[1]; ++j) first += dsm[p + (long long)side[2 * j] * cols + side[2 * j + 1]];
            for (int j = ss[1]; j < ss[2]; ++j) second += dsm[p + (long long)side[2 * j] * cols + side[2 * j + 1]];
            a = 270.0 - angle;
            if (first > second) a -= 180.0;
            if (a < 0.0) a += 360.0;
        }
    }
    if (a == 0.0) {
        // np.gradient with unit-free spacing ``step``: central inside, one-sided on the border
        double fy, fx;
        if (r == 0) fy = (dsm[p + cols] - dsm[p]) / step;
        else if (r == rows - 1) fy = (dsm[p] - dsm[p - cols]) / step;
        else fy = (dsm[p + cols] - dsm[p - cols]) / (2.0 * step);
        if (c == 0) fx = (dsm[p + 1] - dsm[p]) / step;
        else if (c == cols - 1) fx = (dsm[p] - dsm[p - 1]) / step;
        else fx = (dsm[p + 1] - dsm[p - 1]) / (2.0 * step);
        double g = -atan2(fx, fy);
        if (g < 0.0) g += 2.0 * 3.141592653589793;
        a = a + g / (3.141592653589793 / 180.0);
    }
    aspect[p] = a;
}

int launch_adjoint_hour(const sw_domain *dom, const int32_t *bs, int total, int hour, bool top,
                        const sw_adjoint *adj, const int32_t *gofs, const double *gb, int first,
                        int second, double inv_n, hipStream_t s) {
    // the prep stages tiles as the pixel pass does; the gather stages coefficient boxes
    const int cap = lds_records(second);
    const size_t bytes = (size_t)cap * sizeof(float) * 2;
    if (top)
        hipLaunchKernelGGL(k_adj_prep<true>, dim3(total), dim3(BLOCK), bytes, s, dom, bs, total, hour, adj,
                           gofs, gb, first, second, inv_n, cap);
    else
        hipLaunchKernelGGL(k_adj_prep<false>, dim3(total), dim3(BLOCK), bytes, s, dom, bs, total, hour, adj,
                           gofs, gb, first, second, inv_n, cap);
    int rc = hip_check(hipGetLastError(), "k_adj_prep launch");
    if (rc) return rc;
    hipLaunchKernelGGL(k_adj_reduce, dim3(SWA_PARTIAL_SLOTS), dim3(256), 0, s, adj, total, hour,
                       (uint32_t)((1ull << SWA_PARTIAL_SLOTS) - 1));
    rc = hip_check(hipGetLastError(), "k_adj_reduce launch");
    if (rc) return rc;
    if (top) {
        hipLaunchKernelGGL(k_adj_patch, dim3(ADJ_PATCH_BLOCKS, SW_VIS_WORDS), dim3(256), 0, s, dom, adj);
        rc = hip_check(hipGetLastError(), "k_adj_patch launch");
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_adj_gather, dim3(total), dim3(BLOCK), GATHER_LDS_BYTES, s, dom, bs, total, hour, adj, gofs,
                       second, GATHER_LDS_BYTES);
    rc = hip_check(hipGetLastError(), "k_adj_gather launch");
    if (rc) return rc;
    hipLaunchKernelGGL(k_adj_reduce, dim3(SWA_PARTIAL_SLOTS), dim3(256), 0, s, adj, total, hour,
                       (uint32_t)((1u << SWA_TA) | (1u << SWA_PSI) | (1u << SWA_CITGG)));
    return hip_check(hipGetLastError(), "k_adj_reduce launch");
}

// =====================================================================================
// Solar position (replaces Solweig_2015a_metdata_noload + sun_position, umep_solweig.py:2412-2503,
// 2529-3521): the NREL SPA of Reda & Andreas (2003) with delta_t = 0 and the fixed 283 K /
// 1010 mbar refraction, in FP64, with the host restatement's arithmetic order. A forcing row is
// sampled half a timestep before its stamp (seconds dropped), a zenith in (89, 90] becomes 89,
// and altmax is the altitude of the last of the 15-minute samples from 10:15 local before the
// altitude stops increasing (0 when the first sample is below the horizon), found at the first
// row of a chunk and at every midnight row.
#include "spa_terms.inc"

__constant__ int SPA_MONTH_START[13] = {0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334, 365};
constexpr long long SPA_DAY_US = 86400000000LL;

struct SunSite {
    double latitude, longitude, altitude_m, utc;
};

__device__ __forceinline__ bool spa_leap(int y) { return (y % 4 == 0 && y % 100 != 0) || y % 400 == 0; }

__device__ __forceinline__ double spa_fold(double v, double upper) {
    v = v - upper * floor(v / upper);
    return v < 0.0 ? v + upper : v;
}

// sum_k (sum over group g0 + k of A cos(B + C jme)) jme^k / 1e8
__device__ double spa_series(int g0, int ng, double jme) {
    double total = 0.0;
    for (int k = 0; k < ng; ++k) {
        const int s = SPA_GROUP[g0 + k][0], n = SPA_GROUP[g0 + k][1];
        double part = 0.0;
        for (int i = s; i < s + n; ++i) part += SPA_TERMS[i][0] * cos(SPA_TERMS[i][1] + SPA_TERMS[i][2] * jme);
        total = k == 0 ? part : total + part * (k == 1 ? jme : pow(jme, (double)k));
    }
    return total / 1e8;
}

// Julian day of a local calendar instant ``us`` microseconds after 1 January 00:00 of ``year``
// (either side of that year), seconds dropped.
__device__ double spa_julian(int year, long long us, double utc) {
    while (us < 0) {
        year -= 1;
        us += (spa_leap(year) ? 366 : 365) * SPA_DAY_US;
    }
    while (us >= (spa_leap(year) ? 366 : 365) * SPA_DAY_US) {
        us -= (spa_leap(year) ? 366 : 365) * SPA_DAY_US;
        year += 1;
    }
    const int d0 = (int)(us / SPA_DAY_US);
    const long long rem = us % SPA_DAY_US;
    const int hour = (int)(rem / 3600000000LL), minute = (int)((rem / 60000000LL) % 60);
    const int leap = spa_leap(year) ? 1 : 0;
    int month = 1;
    while (month < 12 && d0 >= SPA_MONTH_START[month] + (month >= 2 ? leap : 0)) ++month;
    const int day = d0 - (SPA_MONTH_START[month - 1] + (month - 1 >= 2 ? leap : 0)) + 1;
    const double y = month <= 2 ? year - 1 : year, m = month <= 2 ? month + 12 : month;
    const double ut = ((hour - utc) / 24.0) + (minute / 1440.0) + (0.0 / 86400.0);
    const double d = day + ut;
    double b = 0.0;
    if (!(year < 1582 || (year == 1582 && (month < 10 || (month == 10 && day <= 4))))) {
        const double a = floor(y / 100.0);
        b = 2.0 - a + floor(a / 4.0);
    }
    return d + b + floor(365.25 * (y + 4716.0)) + floor(30.6001 * (m + 1.0)) - 1524.5;
}

// Topocentric zenith and azimuth (degrees) at Julian day jd.
__device__ void spa_point(double jd, const SunSite &site, double &zenith, double &azimuth) {
    const double PI = 3.141592653589793, D = PI / 180.0;      // D: the nutation argument's factor
    const double jc = (jd - 2451545.0) / 36525.0, jce = jc, jme = jc / 10.0;
    const double lon = spa_fold(spa_series(0, 6, jme) * 180.0 / PI, 360.0);
    double b0 = 0.0, b1 = 0.0;
    for (int i = SPA_GROUP[6][0]; i < SPA_GROUP[6][0] + SPA_GROUP[6][1]; ++i)
        b0 += SPA_TERMS[i][0] * cos(SPA_TERMS[i][1] + SPA_TERMS[i][2] * jme);
    for (int i = SPA_GROUP[7][0]; i < SPA_GROUP[7][0] + SPA_GROUP[7][1]; ++i)
        b1 += SPA_TERMS[i][0] * cos(SPA_TERMS[i][1] + SPA_TERMS[i][2] * jme);
    const double lat = spa_fold(((b0 + b1 * jme) / 1e8) * 180.0 / PI, 360.0);
    const double radius = spa_series(8, 5, jme);
    const double geo_lon = spa_fold(lon + 180.0, 360.0), geo_lat = spa_fold(-lat, 360.0);
    const double j2 = pow(jce, 2.0), j3 = pow(jce, 3.0);
    const double x[5] = {
        (1.0 / 189474.0) * j3 + -0.0019142 * j2 + 445267.11148 * jce + 297.85036,
        -(1.0 / 300000.0) * j3 + -0.0001603 * j2 + 35999.05034 * jce + 357.52772,
        (1.0 / 56250.0) * j3 + 0.0086972 * j2 + 477198.867398 * jce + 134.96298,
        (1.0 / 327270.0) * j3 + -0.0036825 * j2 + 483202.017538 * jce + 93.27191,
        (1.0 / 450000.0) * j3 + 0.0020708 * j2 + -1934.136261 * jce + 125.04452};
    double dpsi = 0.0, deps = 0.0;
    for (int i = 0; i < SPA_NUT_ROWS; ++i) {
        double arg = 0.0;
        for (int j = 0; j < 5; ++j) arg += SPA_NUT_ARGS[i][j] * x[j];
        arg *= D;
        dpsi += (SPA_NUT_COEFFS[i][0] + SPA_NUT_COEFFS[i][1] * jce) * sin(arg);
        deps += (SPA_NUT_COEFFS[i][2] + SPA_NUT_COEFFS[i][3] * jce) * cos(arg);
    }
    dpsi /= 36000000.0;
    deps /= 36000000.0;
    const double u = jme / 10.0;
    const double poly[11] = {2.45, 5.79, 27.87, 7.12, -39.05, -249.67, -51.38, 1999.25, -1.55, -4680.93, 84381.448};
    double eps0 = 0.0;
    for (int k = 0; k < 9; ++k) eps0 = eps0 + poly[k] * pow(u, (double)(10 - k));
    eps0 = eps0 + poly[9] * u + poly[10];
    const double eps = eps0 / 3600.0 + deps;
    const double aberration = -20.4898 / (3600.0 * radius);
    const double lam = geo_lon + dpsi + aberration;
    const double nu0 = 280.46061837 + (360.98564736629 * (jd - 2451545.0)) + (0.000387933 * pow(jc, 2.0)) -
                       (pow(jc, 3.0) / 38710000.0);
    const double nu = spa_fold(nu0, 360.0) + dpsi * cos(eps * PI / 180.0);
    const double alpha = spa_fold(atan2(sin(lam * PI / 180.0) * cos(eps * PI / 180.0) - tan(geo_lat * PI / 180.0) * sin(eps * PI / 180.0), cos(lam * PI / 180.0)) *
                                      180.0 / PI, 360.0);
    const double delta = asin(sin(geo_lat * PI / 180.0) * cos(eps * PI / 180.0) + cos(geo_lat * PI / 180.0) * sin(eps * PI / 180.0) * sin(lam * PI / 180.0)) *
                         180.0 / PI;
    const double h = spa_fold(nu + site.longitude - alpha, 360.0);
    const double xi = 8.794 / (3600.0 * radius);
    const double uu = atan(0.99664719 * tan(site.latitude * PI / 180.0));
    const double xx = cos(uu) + (site.altitude_m / 6378140.0) * cos(site.latitude * PI / 180.0);
    const double yy = 0.99664719 * sin(uu) + (site.altitude_m / 6378140.0) * sin(site.latitude * PI / 180.0);
    const double dalpha = atan2(-xx * sin(xi * PI / 180.0) * sin(h * PI / 180.0), cos(delta * PI / 180.0) - xx * sin(xi * PI / 180.0) * cos(h * PI / 180.0));
    const double delta_t = atan2((sin(delta * PI / 180.0) - yy * sin(xi * PI / 180.0)) * cos(dalpha),
                                 cos(delta * PI / 180.0) - yy * sin(xi * PI / 180.0) * cos(h * PI / 180.0)) * 180.0 / PI;
    const double h_t = h - dalpha * 180.0 / PI;
    const double e0 = asin(sin(site.latitude * PI / 180.0) * sin(delta_t * PI / 180.0) +
                           cos(site.latitude * PI / 180.0) * cos(delta_t * PI / 180.0) * cos(h_t * PI / 180.0)) * 180.0 / PI;
    const double refraction = 1.02 / (60.0 * tan((e0 + 10.3 / (e0 + 5.11)) * PI / 180.0));
    zenith = 90.0 - (e0 + refraction);
    azimuth = spa_fold(atan2(sin(h_t * PI / 180.0), cos(h_t * PI / 180.0) * sin(site.latitude * PI / 180.0) -
                                               tan(delta_t * PI / 180.0) * cos(site.latitude * PI / 180.0)) * 180.0 / PI +
                           180.0, 360.0);
}

__device__ __forceinline__ double met_dectime(const double *met, int cols, int i) {
    return met[(long long)i * cols + 1] + met[(long long)i * cols + 2] / 24.0 + met[(long long)i * cols + 3] / (60.0 * 24.0);
}

// The row that starts row i's day: the first row of its chunk or the last midnight row before it.
__device__ __forceinline__ bool sun_day_start(const double *met, int cols, int i, int chunk) {
    if (i % chunk == 0) return true;
    const double dt = met_dectime(met, cols, i);
    return dt == floor(dt);
}

constexpr int ALTMAX_SAMPLES = 128;

// One workgroup per row; rows that start a day search the day's maximum altitude: 15-minute
// samples from 10:15 local, evaluated in parallel, then scanned in order (sequentially past the
// first 128 samples, which only a sun still climbing after 42 hours would need).
__global__ __launch_bounds__(ALTMAX_SAMPLES) void k_sun_altmax(const double *__restrict__ met, int n, int cols,
                                                               int chunk, SunSite site, double *__restrict__ out) {
    const int i = (int)blockIdx.x;
    if (i >= n || !sun_day_start(met, cols, i, chunk)) return;
    __shared__ double zen[ALTMAX_SAMPLES];
    const int year = (int)met[(long long)i * cols], doy = (int)met[(long long)i * cols + 1];
    const long long base = (long long)(doy - 1) * SPA_DAY_US + 600LL * 60000000LL;
    const int k = (int)threadIdx.x + 1;
    double z, a;
    spa_point(spa_julian(year, base + (long long)k * 15LL * 60000000LL, site.utc), site, z, a);
    zen[threadIdx.x] = z;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double sunmax = -90.0, zenith = 90.0;
    for (int s = 1;; ++s) {
        if (!(sunmax <= 90.0 - zenith)) break;
        sunmax = 90.0 - zenith;
        if (s <= ALTMAX_SAMPLES) {
            zenith = zen[s - 1];
        } else {
            spa_point(spa_julian(year, base + (long long)s * 15LL * 60000000LL, site.utc), site, zenith, a);
            if (s > 4 * 96) break;                    // a day has a maximum: never reached in practice
        }
    }
    out[5LL * n + i] = sunmax;
}

// One thread per row: the sun half a timestep before the row's stamp, and the row's day maximum.
__global__ void k_sun_rows(const double *__restrict__ met, int n, int cols, int chunk, SunSite site,
                           double *__restrict__ out) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int c0 = i - i % chunk, len = min(chunk, n - c0);
    const double dectime = met_dectime(met, cols, i);
    const double half = len == 1 ? 0.0 : (met_dectime(met, cols, c0 + 1) - met_dectime(met, cols, c0)) / 2.0;
    const int year = (int)met[(long long)i * cols], doy = (int)met[(long long)i * cols + 1];
    const long long us = (long long)(doy - 1) * SPA_DAY_US + llrint(met[(long long)i * cols + 2] * 3600e6) +
                         llrint(met[(long long)i * cols + 3] * 60e6) - llrint(half * 86400e6);
    double zenith, azimuth;
    spa_point(spa_julian(year, us, site.utc), site, zenith, azimuth);
    if (zenith > 89.0 && zenith <= 90.0) zenith = 89.0;
    out[i] = 90.0 - zenith;
    out[(long long)n + i] = azimuth;
    out[2LL * n + i] = zenith * (3.141592653589793 / 180.0);
    out[3LL * n + i] = (double)doy;
    out[4LL * n + i] = dectime;
    if (!sun_day_start(met, cols, i, chunk)) {
        int j = i;
        while (!sun_day_start(met, cols, j, chunk)) --j;
        out[5LL * n + i] = out[5LL * n + j];
    }
}

}  // namespace

// Scratch of sw_gvf_noalloc: the per-pixel neighbour samples of the ground-view walk, in a
// library-owned device array instead of a caller's work buffer. Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device address): it looks up
// the array's address on each device once (hipGetSymbolAddress, with that device current) and
// passes that pointer.
static __device__ __align__(256) unsigned char scratch[SW_GVF_SCRATCH_BYTES];

namespace {

constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_lock;
void *g_scratch_ptr[SCRATCH_DEVICES] = {};

// The device address of ``scratch`` on the current device (looked up on the first call there).
int scratch_on_current_device(void **out) {
    int dev = -1;
    int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= SCRATCH_DEVICES) return fail(SW_ERR_LIMIT, "sw_gvf_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        if (!p) return fail(SW_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SW_OK;
}

}  // namespace

extern "C" {

int sw_version(void) { return 100; }

int sw_record_bytes(int kind) {
    return kind == 0 ? (int)sizeof(sw_domain) : (kind == 1 ? (int)sizeof(sw_adjoint) : -1);
}

const char *sw_last_error(void) { return g_error.c_str(); }

int sw_device_count(int *count) {
    if (!count) return fail(SW_ERR_ARGUMENT, "count is NULL");
    return hip_check(hipGetDeviceCount(count), "hipGetDeviceCount");
}

int sw_set_device(int device) { return hip_check(hipSetDevice(device), "hipSetDevice"); }

int sw_malloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(SW_ERR_ARGUMENT, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return SW_OK;
    return hip_check(hipMalloc(ptr, bytes), "hipMalloc");
}

int sw_free(void *ptr) { return ptr ? hip_check(hipFree(ptr), "hipFree") : SW_OK; }

int sw_host_alloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(SW_ERR_ARGUMENT, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return SW_OK;
    return hip_check(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
}

int sw_host_free(void *ptr) { return ptr ? hip_check(hipHostFree(ptr), "hipHostFree") : SW_OK; }

int sw_event_sync(void *event) {
    if (!event) return fail(SW_ERR_ARGUMENT, "event is NULL");
    return hip_check(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
}

int sw_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return SW_OK;
    return hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream),
                     "hipMemcpyAsync H2D");
}

int sw_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return SW_OK;
    int rc = hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream),
                       "hipMemcpyAsync D2H");
    if (rc) return rc;
    return hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

int sw_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return SW_OK;
    return hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream),
                     "hipMemcpyAsync D2H");
}

int sw_stream_wait_event(void *stream, void *event) {
    if (!event) return fail(SW_ERR_ARGUMENT, "sw_stream_wait_event: NULL event");
    return hip_check(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0), "hipStreamWaitEvent");
}

int sw_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return SW_OK;
    return hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream),
                     "hipMemcpyAsync D2D");
}

int sw_memset(void *dst, int value, size_t bytes, void *stream) {
    if (!bytes) return SW_OK;
    return hip_check(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream), "hipMemsetAsync");
}

int sw_stream_create(void **stream) {
    if (!stream) return fail(SW_ERR_ARGUMENT, "stream is NULL");
    return hip_check(hipStreamCreateWithFlags((hipStream_t *)stream, hipStreamNonBlocking),
                     "hipStreamCreate");
}

int sw_stream_destroy(void *stream) {
    return stream ? hip_check(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy") : SW_OK;
}

int sw_stream_sync(void *stream) {
    return hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

int sw_event_create(void **event) {
    if (!event) return fail(SW_ERR_ARGUMENT, "event is NULL");
    return hip_check(hipEventCreate((hipEvent_t *)event), "hipEventCreate");
}

int sw_event_destroy(void *event) {
    return event ? hip_check(hipEventDestroy((hipEvent_t)event), "hipEventDestroy") : SW_OK;
}

int sw_event_record(void *event, void *stream) {
    return hip_check(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
}

int sw_event_elapsed_ms(void *start, void *stop, float *ms) {
    int rc = hip_check(hipEventSynchronize((hipEvent_t)stop), "hipEventSynchronize");
    if (rc) return rc;
    return hip_check(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
}

int sw_device_sync(void) { return hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

int sw_svf_visibility(const float *dsm, const float *trees, const float *march_trees, const float *hmax,
                      int hmax_cols, double raise, int rows, int cols, const int32_t *offsets,
                      const double *heights, const int32_t *lengths, int capacity, int patches, uint32_t *bits,
                      void *stream) {
    if (!dsm || !trees || !offsets || !heights || !lengths || !bits)
        return fail(SW_ERR_ARGUMENT, "sw_svf_visibility: NULL pointer");
    if (rows <= 0 || cols <= 0 || capacity <= 0 || patches <= 0 || patches > 32 * SW_VIS_WORDS)
        return fail(SW_ERR_ARGUMENT, "sw_svf_visibility: bad sizes");
    if ((long long)rows * cols > SW_MAX_PIXELS)
        return fail(SW_ERR_LIMIT, "sw_svf_visibility: raster above SW_MAX_PIXELS");
    const int words = (patches + 31) / 32;
    if (hmax) {
        if (capacity % CHUNK != 0 || hmax_cols != (cols + (1 << HCELL_SHIFT) - 1) >> HCELL_SHIFT)
            return fail(SW_ERR_ARGUMENT, "sw_svf_visibility: the march needs capacity % 8 == 0 and "
                                         "hmax_cols == ceil(cols / 16)");
        const int tiles = ((cols + 15) / 16) * ((rows + 15) / 16);
        hipLaunchKernelGGL(k_svf_march, dim3(tiles, words), dim3(BLOCK), 0, (hipStream_t)stream, dsm, trees,
                           march_trees, hmax, hmax_cols, raise, rows, cols, offsets, heights, lengths, capacity,
                           patches, bits, tiles);
        return hip_check(hipGetLastError(), "k_svf_march launch");
    }
    const int tiles = ((cols + TILE_X - 1) / TILE_X) * ((rows + TILE_Y - 1) / TILE_Y);
    hipLaunchKernelGGL(k_svf_visibility, dim3(tiles, words), dim3(BLOCK), 0, (hipStream_t)stream,
                       dsm, trees, raise, rows, cols, offsets, heights, lengths, capacity, patches,
                       bits, tiles);
    return hip_check(hipGetLastError(), "k_svf_visibility launch");
}

int sw_svf_reduce(const uint32_t *bits, const float *dsm, const float *trees, double raise, int rows,
                  int cols, const uint8_t *sectors, const int32_t *band_offsets,
                  const double *weights, int bands, int patches, double trans, double *out,
                  void *stream) {
    if (!bits || !dsm || !trees || !sectors || !band_offsets || !weights || !out)
        return fail(SW_ERR_ARGUMENT, "sw_svf_reduce: NULL pointer");
    if (rows <= 0 || cols <= 0 || bands <= 0 || patches != SW_PATCHES)
        return fail(SW_ERR_ARGUMENT, "sw_svf_reduce: bad sizes");
    const long long P = (long long)rows * cols;
    hipLaunchKernelGGL(k_svf_reduce, dim3((unsigned)((P + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, bits, dsm, trees, raise, P, sectors, band_offsets,
                       weights, bands, trans, out);
    return hip_check(hipGetLastError(), "k_svf_reduce launch");
}

int sw_sky_sums(const uint32_t *bits, long long P, const float *tab, float *out, void *stream) {
    if (!bits || !tab || !out || P <= 0) return fail(SW_ERR_ARGUMENT, "sw_sky_sums: bad arguments");
    hipLaunchKernelGGL(k_sky_sums, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, (hipStream_t)stream, bits,
                       P, tab, out);
    return hip_check(hipGetLastError(), "k_sky_sums launch");
}

int sw_hour_shadow(const sw_domain *doms, const int32_t *bs, int nd, int total, int hour,
                   int full_planes, void *stream) {
    int rc = check_launch(doms, bs, nd, total);
    if (rc) return rc;
    if (hour < 0) return fail(SW_ERR_ARGUMENT, "sw_hour_shadow: negative hour");
    return launch_hour_shadow(doms, bs, nd, total, hour, full_planes != 0, (hipStream_t)stream);
}

int sw_hour_pixels(const sw_domain *doms, const int32_t *bs, int nd, int total, int hour,
                   const int32_t *gofs, const double *gb, int first, int second, int full,
                   void *stream) {
    int rc = check_launch(doms, bs, nd, total);
    if (rc) return rc;
    if (!gofs || !gb) return fail(SW_ERR_ARGUMENT, "sw_hour_pixels: NULL GVF geometry");
    if (first < 1 || second < 1 || second > MAX_SECOND || first > second)
        return fail(SW_ERR_ARGUMENT, "sw_hour_pixels: bad GVF step counts");
    if (hour < 0) return fail(SW_ERR_ARGUMENT, "sw_hour_pixels: negative hour");
    return launch_hour_pixels(doms, bs, nd, total, hour, gofs, gb, first, second, full != 0,
                              (hipStream_t)stream);
}

int sw_hours_forward(const sw_domain *doms, const int32_t *bs, int nd, int total, int h0, int h1,
                     const int32_t *gofs, const double *gb, int first, int second, void *stream) {
    int rc = check_launch(doms, bs, nd, total);
    if (rc) return rc;
    if (!gofs || !gb || first < 1 || second < 1 || second > MAX_SECOND || first > second || h0 < 0 ||
        h1 < h0)
        return fail(SW_ERR_ARGUMENT, "sw_hours_forward: bad arguments");
    for (int h = h0; h < h1; ++h) {
        rc = launch_hour_shadow(doms, bs, nd, total, h, false, (hipStream_t)stream);
        if (rc) return rc;
        rc = launch_hour_pixels(doms, bs, nd, total, h, gofs, gb, first, second, false,
                                (hipStream_t)stream);
        if (rc) return rc;
    }
    return SW_OK;
}

int sw_day_shadow(const sw_domain *doms, const int32_t *bs, int nd, int total, int h0, int h1, void *stream) {
    int rc = check_launch(doms, bs, nd, total);
    if (rc) return rc;
    if (h0 < 0 || h1 < h0 || h1 > 65535) return fail(SW_ERR_ARGUMENT, "sw_day_shadow: bad hour range");
    if (h1 == h0) return SW_OK;
    return launch_hour_shadow(doms, bs, nd, total, h0, false, (hipStream_t)stream, h1 - h0, true);
}

int sw_day_pixels(const sw_domain *doms, const int32_t *bs, int nd, int total, int h0, int h1,
                  const int32_t *gofs, const double *gb, int first, int second, void *stream) {
    int rc = check_launch(doms, bs, nd, total);
    if (rc) return rc;
    if (!gofs || !gb || first < 1 || second < 1 || second > MAX_SECOND || first > second || h0 < 0 || h1 < h0)
        return fail(SW_ERR_ARGUMENT, "sw_day_pixels: bad arguments");
    if (h1 == h0) return SW_OK;
    return launch_day_pixels(doms, bs, nd, total, h0, h1, gofs, gb, first, second, (hipStream_t)stream);
}

int sw_day_pixels_capped(const sw_domain *doms, const int32_t *bs, int nd, int total, int h0, int h1,
                         const int32_t *gofs, const double *gb, int first, int second, int groups_per_cu,
                         void *stream) {
    int rc = check_launch(doms, bs, nd, total);
    if (rc) return rc;
    if (!gofs || !gb || first < 1 || second < 1 || second > MAX_SECOND || first > second || h0 < 0 || h1 < h0 ||
        groups_per_cu < 2 || groups_per_cu > 8)
        return fail(SW_ERR_ARGUMENT, "sw_day_pixels_capped: bad arguments");
    if (h1 == h0) return SW_OK;
    return launch_day_pixels(doms, bs, nd, total, h0, h1, gofs, gb, first, second, (hipStream_t)stream,
                             groups_per_cu);
}

int sw_day_diffuse(const sw_domain *doms, const int32_t *bs, int nd, int total, int n_hours, void *stream) {
    const int rc = check_launch(doms, bs, nd, total);
    if (rc) return rc;
    if (n_hours < 1 || n_hours > 65535) return fail(SW_ERR_ARGUMENT, "sw_day_diffuse: bad hour count");
    hipLaunchKernelGGL(k_day_diffuse, dim3((unsigned)total), dim3(BLOCK), 0, (hipStream_t)stream, doms, bs, nd,
                       total, n_hours);
    return hip_check(hipGetLastError(), "k_day_diffuse launch");
}

int sw_day_forward(const sw_domain *doms, const int32_t *bs, int nd, int total, int h0, int h1,
                   const int32_t *gofs, const double *gb, int first, int second, void *stream) {
    const int rc = sw_day_shadow(doms, bs, nd, total, h0, h1, stream);
    return rc ? rc : sw_day_pixels(doms, bs, nd, total, h0, h1, gofs, gb, first, second, stream);
}

int sw_adjoint_hour(const sw_domain *dom, const int32_t *bs, int total, int hour, int objective_hour,
                    const sw_adjoint *adj, const int32_t *gofs, const double *gb, int first, int second,
                    double inv_n, void *stream) {
    int rc = check_launch(dom, bs, 1, total);
    if (rc) return rc;
    if (!adj || !gofs || !gb) return fail(SW_ERR_ARGUMENT, "sw_adjoint_hour: NULL pointer");
    if (first != 1) return fail(SW_ERR_ARGUMENT, "sw_adjoint_hour: first_steps must be 1");
    if (second < 1 || second > MAX_SECOND || hour < 0 || hour > objective_hour)
        return fail(SW_ERR_ARGUMENT, "sw_adjoint_hour: bad arguments");
    return launch_adjoint_hour(dom, bs, total, hour, hour == objective_hour, adj, gofs, gb, first, second,
                               inv_n, (hipStream_t)stream);
}

int sw_day_adjoint(const sw_domain *dom, const int32_t *bs, int total, int objective_hour, int first_hour,
                   const sw_adjoint *adj, const int32_t *gofs, const double *gb, int first, int second,
                   double inv_n, void *stream) {
    int rc = check_launch(dom, bs, 1, total);
    if (rc) return rc;
    if (!adj || !gofs || !gb || first != 1 || second < 1 || second > MAX_SECOND || first_hour < 0 ||
        first_hour > objective_hour)
        return fail(SW_ERR_ARGUMENT, "sw_day_adjoint: bad arguments");
    const hipStream_t s = (hipStream_t)stream;
    for (int h = 0; h < objective_hour; ++h) {
        rc = launch_hour_shadow(dom, bs, 1, total, h, false, s);
        if (rc) return rc;
        rc = launch_hour_pixels(dom, bs, 1, total, h, gofs, gb, first, second, false, s);
        if (rc) return rc;
    }
    for (int h = objective_hour; h >= first_hour; --h) {
        rc = launch_hour_shadow(dom, bs, 1, total, h, false, s);
        if (rc) return rc;
        rc = launch_adjoint_hour(dom, bs, total, h, h == objective_hour, adj, gofs, gb, first, second,
                                 inv_n, s);
        if (rc) return rc;
    }
    return SW_OK;
}

#define SW_SUB_ENTRY(name, which)                                                                   \
    int name(const void *const *fields, const long long *counts, int n_fields, long long n,          \
             const double *params, int n_params, int real_bytes, void *out_dev, void *stream) {      \
        return run_sub(which, fields, counts, n_fields, n, params, n_params, real_bytes, out_dev, stream); \
    }
SW_SUB_ENTRY(sw_cylindric_wedge, 0)
SW_SUB_ENTRY(sw_wave_delay, 1)
SW_SUB_ENTRY(sw_kup, 2)
SW_SUB_ENTRY(sw_lside, 3)
SW_SUB_ENTRY(sw_surface_temperature, 4)
SW_SUB_ENTRY(sw_assemble_tmrt, 5)
#undef SW_SUB_ENTRY

int sw_step_terms(int op, const void *const *fields, const long long *counts, int n_fields, long long n,
                  const double *params, int n_params, int real_bytes, void *out, void *stream) {
    if (!fields || !counts || !params || !out) return fail(SW_ERR_ARGUMENT, "sw_step_terms: NULL pointer");
    if (op < 0 || op >= TERMS_OPS) return fail(SW_ERR_ARGUMENT, "sw_step_terms: op is 0..6");
    if (n_fields != TERMS_NFIELDS[op] || n_params != TERMS_NPARAMS[op])
        return fail(SW_ERR_ARGUMENT, "sw_step_terms: wrong number of fields or parameters");
    if (real_bytes != 4 && real_bytes != 8) return fail(SW_ERR_ARGUMENT, "sw_step_terms: precision is 4 or 8 bytes");
    if (n <= 0) return SW_OK;
    FieldSet F = {};
    const int rc = make_fieldset(fields, counts, n_fields, n, F);
    if (rc) return rc;
    Coef7 q = {};
    for (int k = 0; k < n_params; ++k) q.c[k] = params[k];
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (real_bytes == 8)
        hipLaunchKernelGGL(k_step_terms<double>, grid, block, 0, (hipStream_t)stream, op, F, n, q, (double *)out);
    else
        hipLaunchKernelGGL(k_step_terms<float>, grid, block, 0, (hipStream_t)stream, op, F, n, q, (float *)out);
    return hip_check(hipGetLastError(), "k_step_terms launch");
}

int sw_diffuse(const void *diffsh, const double *lum, long long n, int patches, double radD, int real_bytes, void *out,
               void *stream) {
    if (!diffsh || !lum || !out) return fail(SW_ERR_ARGUMENT, "sw_diffuse: NULL pointer");
    if (real_bytes != 4 && real_bytes != 8) return fail(SW_ERR_ARGUMENT, "sw_diffuse: precision is 4 or 8 bytes");
    if (patches <= 0) return fail(SW_ERR_ARGUMENT, "sw_diffuse: no patches");
    if (n <= 0) return SW_OK;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (real_bytes == 8)
        hipLaunchKernelGGL(k_diffuse<double>, grid, block, 0, (hipStream_t)stream, (const double *)diffsh, lum, n,
                           patches, radD, (double *)out);
    else
        hipLaunchKernelGGL(k_diffuse<float>, grid, block, 0, (hipStream_t)stream, (const float *)diffsh, lum, n,
                           patches, radD, (float *)out);
    return hip_check(hipGetLastError(), "k_diffuse launch");
}

int sw_gvf(const void *const *fields, const long long *counts, int n_fields, int rows, int cols,
           const int32_t *offsets, const double *bounds, int stride, int first, int second,
           const double *params, int n_params, int real_bytes, void *work, void *out, void *stream) {
    if (!fields || !counts || !offsets || !bounds || !params || !work || !out)
        return fail(SW_ERR_ARGUMENT, "sw_gvf: NULL pointer");
    if (n_fields != 10 || n_params != 6) return fail(SW_ERR_ARGUMENT, "sw_gvf: expects 10 fields, 6 parameters");
    if (real_bytes != 4 && real_bytes != 8) return fail(SW_ERR_ARGUMENT, "sw_gvf: precision is 4 or 8 bytes");
    if (rows <= 0 || cols <= 0 || first < 1 || second < 1 || stride < second)
        return fail(SW_ERR_ARGUMENT, "sw_gvf: bad raster or step counts");
    const long long n = (long long)rows * cols;
    FieldSet F = {};
    const int rc = make_fieldset(fields, counts, n_fields, n, F);
    if (rc) return rc;
    const hipStream_t s = (hipStream_t)stream;
    return real_bytes == 8 ? launch_gvf<double>(F, rows, cols, offsets, stride, bounds, first, second, params, work, out, s)
                           : launch_gvf<float>(F, rows, cols, offsets, stride, bounds, first, second, params, work, out, s);
}

int sw_gvf_noalloc(const void *const *fields, const long long *counts, int n_fields, int rows, int cols,
                   const int32_t *offsets, const double *bounds, int stride, int first, int second,
                   const double *params, int n_params, int real_bytes, void *out, void *stream) {
    if (!fields || !counts || !offsets || !bounds || !params || !out)
        return fail(SW_ERR_ARGUMENT, "sw_gvf_noalloc: NULL pointer");
    if (real_bytes != 4 && real_bytes != 8) return fail(SW_ERR_ARGUMENT, "sw_gvf_noalloc: precision is 4 or 8 bytes");
    if (rows <= 0 || cols <= 0) return fail(SW_ERR_ARGUMENT, "sw_gvf_noalloc: bad raster");
    // the samples need rows * cols * 8 values of the working precision (GvfSample)
    const unsigned long long need = (unsigned long long)rows * (unsigned long long)cols * 8ull * (unsigned)real_bytes;
    if (need > SW_GVF_SCRATCH_BYTES)
        return fail(SW_ERR_LIMIT, "sw_gvf_noalloc: " + std::to_string(rows) + " x " + std::to_string(cols) +
                                      " pixels need " + std::to_string(need) + " bytes of samples; the library's " +
                                      "scratch holds " + std::to_string(SW_GVF_SCRATCH_BYTES) +
                                      " (use sw_gvf with a work buffer)");
    static_assert(sizeof(GvfSample<float>) == 8 * 4 && sizeof(GvfSample<double>) == 8 * 8, "GvfSample is 8 values");
    void *work = nullptr;
    const int rc = scratch_on_current_device(&work);
    if (rc) return rc;
    return sw_gvf(fields, counts, n_fields, rows, cols, offsets, bounds, stride, first, second, params, n_params,
                  real_bytes, work, out, stream);
}

int sw_pack_visibility(const void *sh, const void *veg, const void *vb, int mask_bytes, long long n, int patches,
                       uint32_t *flags, void *stream) {
    if (!sh || !veg || !vb || !flags) return fail(SW_ERR_ARGUMENT, "sw_pack_visibility: NULL pointer");
    if (mask_bytes != 1 && mask_bytes != 4 && mask_bytes != 8)
        return fail(SW_ERR_ARGUMENT, "sw_pack_visibility: mask values are 1, 4 or 8 bytes");
    if (n <= 0 || patches <= 0) return fail(SW_ERR_ARGUMENT, "sw_pack_visibility: empty raster or patch set");
    return launch_pack(sh, veg, vb, mask_bytes, n, patches, flags, (hipStream_t)stream);
}

int sw_longwave(const uint32_t *flags, int patches, long long n, const double *table,
                const void *const *fields, const long long *counts, int n_fields, const double *params,
                int n_params, int real_bytes, void *out, void *stream) {
    if (!flags || !table || !fields || !counts || !params || !out)
        return fail(SW_ERR_ARGUMENT, "sw_longwave: NULL pointer");
    if (n_fields != 3 || n_params != 3) return fail(SW_ERR_ARGUMENT, "sw_longwave: expects 3 fields, 3 parameters");
    if (real_bytes != 4 && real_bytes != 8) return fail(SW_ERR_ARGUMENT, "sw_longwave: precision is 4 or 8 bytes");
    if (patches <= 0) return fail(SW_ERR_ARGUMENT, "sw_longwave: no patches");
    if (n <= 0) return SW_OK;
    FieldSet F = {};
    const int rc = make_fieldset(fields, counts, n_fields, n, F);
    if (rc) return rc;
    const hipStream_t s = (hipStream_t)stream;
    return real_bytes == 8 ? launch_longwave<double>(flags, patches, n, table, F, params, out, s)
                           : launch_longwave<float>(flags, patches, n, table, F, params, out, s);
}

int sw_kside(const uint32_t *flags, const void *diffsh, int patches, long long n, const double *table,
             const void *const *fields, const long long *counts, int n_fields, const double *params, int n_params,
             int real_bytes, void *out, void *stream) {
    if (!fields || !counts || !params || !out) return fail(SW_ERR_ARGUMENT, "sw_kside: NULL pointer");
    if (n_fields != 15 || n_params != 16) return fail(SW_ERR_ARGUMENT, "sw_kside: expects 15 fields, 16 parameters");
    if (real_bytes != 4 && real_bytes != 8) return fail(SW_ERR_ARGUMENT, "sw_kside: precision is 4 or 8 bytes");
    KsideParams q = {};
    q.radI = params[0]; q.radD = params[1]; q.radG = params[2]; q.cos_alt = params[3];
    q.one_minus_psi = 1.0 - params[4]; q.albedo = params[5];
    for (int f = 0; f < 4; ++f) { q.sin_face[f] = params[6 + f]; q.lit[f] = params[10 + f] != 0.0; }
    q.cyl = (int)params[14]; q.ani = (int)params[15];
    if (q.ani == 1 && (!flags || !diffsh || !table || patches <= 0))
        return fail(SW_ERR_ARGUMENT, "sw_kside: the anisotropic form needs flags, diffsh and a patch table");
    if (q.ani != 1) patches = 0;
    if (n <= 0) return SW_OK;
    FieldSet F = {};
    const int rc = make_fieldset(fields, counts, n_fields, n, F);
    if (rc) return rc;
    const hipStream_t s = (hipStream_t)stream;
    return real_bytes == 8 ? launch_kside<double>(flags, diffsh, patches, n, table, F, q, out, s)
                           : launch_kside<float>(flags, diffsh, patches, n, table, F, q, out, s);
}

int sw_dsm_walls(const double *dsm, int rows, int cols, double threshold, double step, int radius, int n_angles,
                 const int32_t *line, const int32_t *line_start, const int32_t *side, const int32_t *side_start,
                 uint8_t *occupancy, double *walls, double *aspect, void *stream) {
    if (!dsm || !occupancy || !walls || !aspect) return fail(SW_ERR_ARGUMENT, "sw_dsm_walls: NULL pointer");
    if (rows < 2 || cols < 2 || !(step > 0.0)) return fail(SW_ERR_ARGUMENT, "sw_dsm_walls: raster below 2 x 2 or bad step");
    if (n_angles != 0 && (n_angles > 180 || !line || !line_start || !side || !side_start || radius < 0 ||
                          rows <= 2 * radius + 2 || cols <= 2 * radius + 2))
        return fail(SW_ERR_ARGUMENT, "sw_dsm_walls: bad template tables for this raster");
    const long long n = (long long)rows * cols;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    const hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_wall_heights, grid, block, 0, s, dsm, rows, cols, threshold, walls, occupancy);
    hipLaunchKernelGGL(k_wall_aspect, grid, block, 0, s, dsm, walls, occupancy, rows, cols, radius, n_angles, line,
                       line_start, side, side_start, step, aspect);
    return hip_check(hipGetLastError(), "wall kernels launch");
}

int sw_gvf_static(const uint8_t *ground, const float *alb, const float *walls, const float *aspect, int rows,
                  int cols, const int32_t *gofs, const double *gbounds, int first, int second, uint32_t *len,
                  float *nosh, uint32_t *edge, void *stream) {
    if (!ground || !alb || !walls || !aspect || !gofs || !gbounds || !len || !nosh)
        return fail(SW_ERR_ARGUMENT, "sw_gvf_static: NULL pointer");
    if (rows <= 0 || cols <= 0 || first < 1 || second < 1 || second > MAX_SECOND)
        return fail(SW_ERR_ARGUMENT, "sw_gvf_static: bad raster or step counts");
    const long long n = (long long)rows * cols;
    hipLaunchKernelGGL(k_gvf_static, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, ground,
                       alb, walls, aspect, rows, cols, gofs, gbounds, first, second, len, nosh, edge);
    return hip_check(hipGetLastError(), "k_gvf_static launch");
}

int sw_gvf_fixed_steps(int32_t *out, int n) {
    if (!out) return fail(SW_ERR_ARGUMENT, "sw_gvf_fixed_steps: NULL pointer");
    if (n < 0 || n > GV_FIXED_STEPS) return fail(SW_ERR_ARGUMENT, "sw_gvf_fixed_steps: n must be 0..22");
    for (int d = 0; d < 18; ++d)
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < 2; ++j) out[(d * n + k) * 2 + j] = GV_STEPS[d][k][j];
    return SW_OK;
}

int sw_utci_scalar(const float *tmrt, long long n, const double *coefficients, double ta32,
                   int sentinel, const uint8_t *ground, float *out, void *stream) {
    if (!tmrt || !out || !coefficients) return fail(SW_ERR_ARGUMENT, "sw_utci_scalar: NULL pointer");
    if (n <= 0) return SW_OK;
    Coef7 cf;
    for (int k = 0; k < 7; ++k) cf.c[k] = coefficients[k];
    const long long blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_utci_scalar, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0,
                       (hipStream_t)stream, tmrt, n, cf, (float)ta32, sentinel, ground, out);
    return hip_check(hipGetLastError(), "k_utci_scalar launch");
}

int sw_sun_position(const double *met, int n_rows, int met_cols, int chunk_rows, double latitude, double longitude,
                    double altitude_m, double utc, double *out, void *stream) {
    if (!met || !out || n_rows < 0 || met_cols < 4 || chunk_rows < 0)
        return fail(SW_ERR_ARGUMENT, "sw_sun_position: bad arguments");
    if (n_rows == 0) return SW_OK;
    const int chunk = chunk_rows == 0 ? n_rows : chunk_rows;
    const SunSite site = {latitude, longitude, altitude_m, utc};
    const hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_sun_altmax, dim3(n_rows), dim3(ALTMAX_SAMPLES), 0, s, met, n_rows, met_cols, chunk, site, out);
    int rc = hip_check(hipGetLastError(), "k_sun_altmax launch");
    if (rc) return rc;
    hipLaunchKernelGGL(k_sun_rows, dim3((n_rows + 63) / 64), dim3(64), 0, s, met, n_rows, met_cols, chunk, site, out);
    return hip_check(hipGetLastError(), "k_sun_rows launch");
}

int sw_copy(void *dst, const void *src, size_t bytes, void *stream) {
    if (!dst || !src) return fail(SW_ERR_ARGUMENT, "sw_copy: NULL pointer");
    if (bytes % 16 != 0 || ((uintptr_t)dst | (uintptr_t)src) % 16 != 0)
        return fail(SW_ERR_ARGUMENT, "sw_copy: size and addresses must be 16-byte multiples");
    const long long n = (long long)(bytes / 16);
    hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, (hipStream_t)stream, (const float4 *)src,
                       (float4 *)dst, n);
    return hip_check(hipGetLastError(), "k_copy launch");
}

int sw_graph_begin(void *stream) {
    return hip_check(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal),
                     "hipStreamBeginCapture");
}

int sw_graph_end(void *stream, void **graph_exec) {
    if (!graph_exec) return fail(SW_ERR_ARGUMENT, "graph_exec is NULL");
    hipGraph_t g = nullptr;
    int rc = hip_check(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture");
    if (rc) return rc;
    hipGraphExec_t e = nullptr;
    rc = hip_check(hipGraphInstantiate(&e, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    (void)hipGraphDestroy(g);
    if (rc) return rc;
    *graph_exec = (void *)e;
    return SW_OK;
}

int sw_graph_launch(void *graph_exec, void *stream) {
    return hip_check(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream), "hipGraphLaunch");
}

int sw_graph_destroy(void *graph_exec) {
    return graph_exec ? hip_check(hipGraphExecDestroy((hipGraphExec_t)graph_exec), "hipGraphExecDestroy")
                      : SW_OK;
}

}  // extern "C"
