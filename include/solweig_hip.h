/* C ABI of libsolweig_hip.so - the MI355X (gfx950) SOLWEIG 2022a radiation engine.
 *
 * Every function returns 0 on success and a negative code on failure; sw_last_error() gives
 * the message of the last failure on the calling thread. Pointers named *_dev are device
 * pointers (from sw_malloc or any HIP allocation); ``stream`` is a hipStream_t (NULL = the
 * null stream). Rasters are row-major [rows][cols]; planes are stacked [plane][rows*cols].
 *
 * Each entry point names the reference interface it replaces (zetingli/SOLWEIG-CUDA).
 */
#ifndef SOLWEIG_HIP_H
#define SOLWEIG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SW_OK = 0,
    SW_ERR_ARGUMENT = -1,
    SW_ERR_HIP = -2,
    SW_ERR_LIMIT = -3,
};

/* Largest raster (pixels) of one domain: the ray marches address the height planes with 32-bit
 * byte offsets.                                                                               */
#define SW_MAX_PIXELS ((1LL << 30) - 1)

/* ----- per-hour parameter record: SW_HOUR_STRIDE doubles per (domain, hour) ------------- */
enum {
    SWH_DAY = 0,          /* 1 when the sun altitude > 0 (model.py:96)                        */
    SWH_ALTITUDE, SWH_AZIMUTH, SWH_TANZEN,
    SWH_TA, SWH_RADG, SWH_RADD, SWH_RADI, SWH_PSI,
    SWH_SINALT, SWH_COSALT,                  /* sin/cos(altitude*pi/180)                      */
    SWH_ALTMAX, SWH_FRAC, SWH_SNUP, SWH_CITGG, /* surface temperature wave (model.py:141-142)  */
    SWH_LAG_WEIGHT, SWH_LAG_FIRST, SWH_LAG_UPDATE, /* TsWaveDelay of the Lup maps (742-756)  */
    SWH_FACE_LOW, SWH_FACE_HIGH, SWH_FACE_MIDDLE,  /* self-shadowed wall bounds (678-684)    */
    SWH_RAY_START, SWH_RAY_LEN,              /* sun ray slice in the ray tables: the start is a
                                                multiple of 8 steps and the slice is padded to a
                                                multiple of 8 (the last step repeated)          */
    SWH_LWALL,            /* SBC*ewall*((Tgwall+Ta+273.15)^4-(Ta+273.15)^4), gvf.cu field 7 */
    SWH_FLUX4,            /* SBC*(Ta+273.15)^4                                             */
    SWH_LW_AMBIENT, SWH_LW_SUNLIT, SWH_LW_REFL,  /* ewall*SBC*T^4/pi, (1-ewall)*0.5/pi     */
    SWH_SW_SUNLIT, SWH_SW_SHADED,            /* (alb*radI*cos+radD/2)/pi, alb*radD/2/pi        */
    SWH_TA32,             /* float32(Ta) as used by the UTCI path (series.py:466)           */
    SWH_UTCI0,            /* 7 coefficients of UTCI in powers of dTmrt                     */
    SWH_PATCH_TABLE = SWH_UTCI0 + 7,         /* row of the patch tables for this hour         */
    SWH_TGOUT_WEIGHT, SWH_TGOUT_UPDATE,      /* TsWaveDelay of TgOut (model.py:161-163)       */
    SWH_ALBEDO_B, SWH_ABSK, SWH_ABSL, SWH_FUP, SWH_FSIDE, SWH_FCYL,
    SWH_DIFF_VEG,         /* diffsh of a vegetation-blocked patch: 1-(1-0.03f) in FP32       */
    SWH_TGWALL,
    SWH_SKY_A, SWH_SKY_B, /* sky patch flux = A*omega*trig - B*g*omega*trig, A = SBC*T^4/pi,
                             B = (1 - esky)*A, g = exp(0.308*(1.7 - 1/cos(zenith))) per band */
    SWH_ELIG,             /* 5 values: the 32-patch words of the sun-test eligibility bits
                             (patch k: bit k % 32 of word k / 32), as exact integers          */
    SWH_TG_SINE = SWH_ELIG + 5, /* sin(((frac - SNUP/24) / (TmaxLST/24 - SNUP/24)) pi/2) of the
                             surface temperature wave (umep_solweig.py:144-170), host FP64, for
                             the TmaxLST of SWH_TG_TMAX; other TmaxLST values take it in FP32 */
    SWH_TG_TMAX,
    SW_HOUR_STRIDE = 64
};

/* patch tables per hour row: SW_PATCH_F floats per patch ([153][SW_PATCH_F]) and SW_PATCH_D
 * rows of 153 doubles ([SW_PATCH_D][153]: yi of every patch, then the two thresholds, equal
 * within an altitude band)                                                                   */
enum {
    SWP_LV = 0, SWP_KD, SWP_LDN, SWP_LSD, SWP_WCOS, SWP_WSIN, SWP_DE, SWP_DW, SWP_DN, SWP_DS,
    SWP_ELIGIBLE, SW_PATCH_F = 12
};
enum { SWP_YI = 0, SWP_SUN_BELOW, SWP_SHADE_FROM, SW_PATCH_D = 4 };
enum { SW_PATCHES = 153, SW_VIS_WORDS = 5 };

/* Per hour row, the 32-bit words of the patch search tables (sw_domain.ptab_s):
 *   SWS_NIB   float [39][16][2]: over the 4 patches 4i..4i+3, the sums of (luminance,
 *             lumChi cos omega) of the patches whose bit is set in the 4-bit index;
 *   SWS_CSUN  double [8][32]: per altitude band, in ascending yi order of its patches, the
 *             breakpoint c_j = the smallest double x >= 0 with x + yi_j (FP64) >= the band's
 *             sunlit threshold: the patch is sunlit from a pixel exactly when tan(asvf) < c_j
 *             (0 after the band's last patch);
 *   SWS_CSH   double [8][32]: the same for the band's shaded threshold (shaded iff
 *             tan(asvf) >= c_j);
 *   SWS_MASK  uint32 [8][33]: per band, the band-local bits (bit i: patch BAND_START + i) of the
 *             j smallest-yi patches, j = 0..32;
 *   SWS_ELIG  uint32 [8]: per band, the band-local sun-test eligibility bits.              */
enum { SWS_NIB = 0, SWS_CSUN = 1248, SWS_CSH = 1760, SWS_MASK = 2272, SWS_ELIG = 2536,
       SW_SEARCH_WORDS = 2560 };

/* The day's diffuse table (sw_domain.diffuse_tab, read by sw_day_diffuse): per hour the patch
 * luminance (Perez, umep_solweig.py:148-154) and lumChi cos(altitude) omega (the diffuse Kside
 * term, shortwave.cu:155-157) as integer rows of a product with the pixels' visibility masks.
 *   rows:    SW_DIFF_ROWS(n_hours), two per hour: 2h luminance, 2h + 1 lumChi cos omega;
 *   columns: SW_DIFF_K patch positions: k < 160 patch k of the sky mask (vsh & vveg), 160 + k
 *            patch k of the vegetation-only mask (vsh & ~vveg), its value times the hour's
 *            diffsh of vegetation (SWH_DIFF_VEG in FP32, the product in FP64); positions
 *            153..159 and 313..319 zero;
 *   layout:  the int8 digits in the order the device's matrix operands take them,
 *            [rows / 16][SW_DIFF_K / 64][SW_DIFF_DIGITS][64][16]: per block of 16 rows, step of
 *            64 positions and digit, 64 fragments of 16 bytes, fragment c + 16 g holding row
 *            16 block + c at positions 64 step + 16 g .. + 15; then double scale [rows].
 * Row r holds integers q (|q| <= 2^27) with value = q * scale_r exactly, scale_r a power of two
 * (the row's values rounded to 27 bits of its largest magnitude), as the digits
 * q = d3 2^21 + d2 2^14 + d1 2^7 + d0, d0..d2 in [0, 127], d3 in [-64, 64]. A negative scale
 * marks the rows of a night hour (and the rows past the day): no sums are written for them.  */
enum { SW_DIFF_K = 320, SW_DIFF_DIGITS = 4 };
#define SW_DIFF_ROWS(n_hours) (16 * (((n_hours) + 7) / 8))
#define SW_DIFF_TABLE_BYTES(n_hours) \
    ((size_t)SW_DIFF_ROWS(n_hours) * (SW_DIFF_DIGITS * SW_DIFF_K + sizeof(double)))

/* FULL-mode output planes of sw_hour_pixels (the per-step 39-tuple's per-pixel members) */
enum {
    SWF_KDOWN = 0, SWF_KUP, SWF_LDOWN, SWF_LUP, SWF_TG, SWF_KEAST, SWF_KSOUTH, SWF_KWEST,
    SWF_KNORTH, SWF_LEAST, SWF_LSOUTH, SWF_LWEST, SWF_LNORTH, SWF_KSIDEI, SWF_TGOUT, SWF_LSIDE,
    SWF_KSIDED, SWF_DRAD, SWF_KSIDE, SWF_SHADOW, SWF_WALLSUN, SWF_FSH, SW_FULL_PLANES
};

/* Per-hour extra outputs of the batched (non-FULL) pixel pass: with sw_domain.full set, hour h
 * of a day writes plane k at full + (h * SW_EXTRA_PLANES + k) * P (the series outputs Kdown,
 * Kup, Ldown, Lup and Shadow of series.py:458-460).                                          */
enum { SWE_KDOWN = 0, SWE_KUP, SWE_LDOWN, SWE_LUP, SWE_SHADOW, SW_EXTRA_PLANES };

/* Device-side description of one domain (320 bytes: 40 fields, all 8 bytes wide); pointers are
 * device pointers. Optional outputs may be NULL. */
typedef struct sw_domain {
    int64_t rows, cols, pixels, landcover;
    double dsm_raise;
    int64_t tile_cols;    /* workgroup tile: tile_cols x (256 / tile_cols) pixels, tile_cols one of
                             16, 32, 64, 128 (0 means 32); block_start counts these tiles       */
    const float *dsm, *trees, *walls, *aspect;               /* [P]; trees clipped >= 0       */
    const float *emis, *alb, *tgk, *tstart, *tmaxlst, *svfbuveg; /* [P]                      */
    const double *tan_asvf, *tan_svfalfa;                    /* [P]                           */
    const uint8_t *ground;                                   /* [P] 1 where dsm-dem < 2        */
    const uint32_t *vis;                                     /* [3][5][P] sh, veg, vb bits     */
    void *rec;            /* [P] 8-byte hour records written by the shadow pass ([hours][P] for
                             sw_day_*: one slot per hour): the Lup field of sunonsurface_2018a,
                             then albedo x shadow whose sign bit marks a sunlit wall (gvf.cu
                             fields 3-5)                                                      */
    float *lag;                                              /* [6][P] Tgmap1, E, S, W, N, TgOut1 */
    float *tmrt, *utci;                                      /* [n_hours][P] (may be NULL)     */
    const double *hours;                                     /* [n_hours][SW_HOUR_STRIDE]     */
    const float *ptab_f;                                     /* [rows][153][SW_PATCH_F]       */
    const double *ptab_d;                                    /* [rows][SW_PATCH_D][153]       */
    const int32_t *ray_off;                                  /* [sum L][2] (row, col) shifts  */
    const double *ray_dz;                                    /* [sum L] ray heights            */
    uint8_t *masks;       /* [P] ([hours][P] for sw_day_*) bits of the shadow pass (read by the
                             pixel pass): 1 not
                             building-shadowed, 2 no canopy shadow, 4 vbshvegsh, 8 sunlit wall,
                             16 facesh                                                         */
    double *shadow_planes;                                   /* [8][P] cast_shadow planes      */
    float *full;          /* FULL mode: [SW_FULL_PLANES][P]; batched mode (optional):
                             [hours][SW_EXTRA_PLANES][P] per-hour extra outputs            */
    const float *skysums;                                    /* [SW_SKY_SUMS][P] static sums   */
    const float *hmax;    /* [3][3][ceil(rows/16)][hmax_cols]: per 16 x 16 cell an upper bound
                             of max(dsm, canopy top) (FP32 rounded up); level (i, j) holds at
                             cell (r, c) its maximum over cells [r, r + 2^i) x [c, c + 2^j).
                             Lets the sun-ray march skip steps                              */
    int64_t hmax_cols;    /* ceil(cols/16)                                                    */
    const uint32_t *gvf_len; /* [5][P] ground-view steps before the first non-ground sample,
                                one byte per direction (direction d: byte d % 4 of word d / 4;
                                static, sw_gvf_static)                                         */
    const float *gvf_nosh;   /* [10][P] static albedo sums of the ground view: a (all, E, S, W,
                                N) then b; gvfalbnosh sum = a + albedo_b * b                    */
    const float *march_trees; /* [P] or NULL: the canopy heights the sun-ray march reads, trees
                                 where (double)trees + a != a (a = dsm + dsm_raise) and -inf where
                                 the sum equals a (no canopy top). Only valid when no pixel has
                                 0.25 * trees + a == a != trees + a; NULL: the march forms the
                                 canopy tops from trees (series.py:252-287)                   */
    const uint32_t *ptab_s;   /* [rows][SW_SEARCH_WORDS] patch search tables, or NULL (the
                                 pixel pass then tests every patch)                            */
    const uint32_t *gvf_edge; /* [5][P] or NULL: per direction (byte d % 4 of word d / 4) the
                                 number of ground-view steps whose shift stays inside the raster
                                 (static, sw_gvf_static); later steps repeat the last of them */
    const void *diffuse_tab;  /* the day's diffuse table (SW_DIFF_K) or NULL                   */
    float *diffuse;           /* [hours][P][2] or NULL: per pixel-hour the diffuse sums
                                 (luminance, lumChi cos omega) over the pixel's sky patches plus
                                 diffsh x those over its vegetation-only patches, written by
                                 sw_day_diffuse; when set, the batched pixel passes read hour h's
                                 sums from slot h instead of summing the patches              */
} sw_domain;

/* Static per-pixel sums over the patches of each visibility class (day patch geometry), the
 * hour-independent part of the longwave patch sums: */
enum {
    SWS_SKY_SIN = 0, SWS_SKY_GSIN, SWS_SKY_COS, SWS_SKY_GCOS,   /* sky patches: w sin, g w sin, ... */
    SWS_VEG_COS, SWS_VEG_SIN,                                 /* vegetation patches              */
    SWS_OBS_COS, SWS_OBS_SIN,                                 /* any obstruction (reflection)    */
    SWS_BLD_COS, SWS_BLD_SIN,                                 /* building patches                */
    SW_SKY_SUMS = 10
};

/* ----- runtime ------------------------------------------------------------------------ */
int sw_version(void);
/* sizeof(sw_domain) (kind 0) or sizeof(sw_adjoint) (kind 1) as this library was built, so that a
 * binding can refuse a library whose records do not match its own; -1 for another kind.
 * No counterpart in the reference (its records are CuPy arrays).                         */
int sw_record_bytes(int kind);
const char *sw_last_error(void);
int sw_device_count(int *count);
int sw_set_device(int device);
int sw_malloc(void **ptr_dev, size_t bytes);
int sw_free(void *ptr_dev);
/* Page-locked host memory (hipHostMalloc) for asynchronous, stream-ordered uploads. */
int sw_host_alloc(void **ptr, size_t bytes);
int sw_host_free(void *ptr);
int sw_memcpy_h2d(void *dst_dev, const void *src, size_t bytes, void *stream);
int sw_memcpy_d2h(void *dst, const void *src_dev, size_t bytes, void *stream);
int sw_memcpy_d2d(void *dst_dev, const void *src_dev, size_t bytes, void *stream);
/* Output path (replaces the side copy stream and pinned transfers of runtime.py:79-163): an
 * asynchronous device-to-host copy into page-locked memory from sw_host_alloc, and a stream
 * wait on an event recorded on another stream.                                            */
int sw_memcpy_d2h_async(void *dst_pinned, const void *src_dev, size_t bytes, void *stream);
int sw_stream_wait_event(void *stream, void *event);
int sw_memset(void *dst_dev, int value, size_t bytes, void *stream);
int sw_stream_create(void **stream);
int sw_stream_destroy(void *stream);
int sw_stream_sync(void *stream);
int sw_event_create(void **event);
int sw_event_destroy(void *event);
int sw_event_record(void *event, void *stream);
int sw_event_elapsed_ms(void *start, void *stop, float *ms);
int sw_event_sync(void *event);
int sw_device_sync(void);

/* ----- static sky view (replaces svf.svf_calculator, svf.py:81-114; geometry.cu:77-159) -- */
/* Visibility bits of 153 patches for every pixel: bit k of word w of mask m is patch 32w+k;
 * m = 0 sky not blocked by buildings (shmat), 1 not vegetation-blocked (vegmat), 2 vbmat.
 * Patch k's ray (svf.py:19-78, from step 0): offsets_dev [patches][capacity][2] (row, col),
 * heights_dev [patches][capacity] FP64, lengths_dev [patches] steps. With hmax_dev (the
 * [3][3][R][C] block maxima of the 16 x 16-cell height bound of max(dsm + raise, canopy top),
 * R = ceil(rows / 16), C = hmax_cols = ceil(cols / 16)) each patch is marched as the hour's sun
 * ray is, with windows below every live lane's height skipped; this needs dsm + raise >= 0
 * everywhere and capacity % 8 == 0. march_trees_dev (may be NULL) is the march's canopy plane
 * (trees, -inf where the canopy top equals the DSM). hmax_dev NULL: one thread per (pixel,
 * 32 patches), any heights.                                                               */
int sw_svf_visibility(const float *dsm_dev, const float *trees_dev, const float *march_trees_dev,
                      const float *hmax_dev, int hmax_cols, double dsm_raise, int rows, int cols,
                      const int32_t *offsets_dev, const double *heights_dev, const int32_t *lengths_dev,
                      int capacity, int patches, uint32_t *bits_dev, void *stream);
/* Band-weighted reduction to 15 SVF planes + svftotal, FP64 [16][P] in the kernel order
 * (total, E, S, W, N for building, vegetation, vegetation-blocking-building).              */
int sw_svf_reduce(const uint32_t *bits_dev, const float *dsm_dev, const float *trees_dev,
                  double dsm_raise, int rows, int cols, const uint8_t *sectors_dev,
                  const int32_t *band_offsets_dev, const double *weights_dev, int bands,
                  int patches, double transmissivity, double *out_dev, void *stream);

/* Static patch sums per pixel from the visibility bits: table_dev [153][4] floats per patch
 * (w sin, w cos, g w sin, g w cos; w the steradian) -> out_dev [SW_SKY_SUMS][P].            */
int sw_sky_sums(const uint32_t *bits_dev, long long pixels, const float *table_dev, float *out_dev,
                void *stream);

/* ----- per-hour path ----------------------------------------------------------------- */
/* Sun-ray shadow march for hour ``hour`` of every domain (replaces model.cast_shadow,
 * model.py:689-708, and geometry.cu:185-253). ``domains_dev`` is a device array of n_domains
 * sw_domain; ``block_start_dev`` [n_domains+1] the cumulative 32x8 tile counts.
 * full_planes != 0 also writes the 8 FP64 cast_shadow planes to shadow_planes.             */
int sw_hour_shadow(const sw_domain *domains_dev, const int32_t *block_start_dev, int n_domains,
                   int total_blocks, int hour, int full_planes, void *stream);
/* Ground view factors, lag recursion, patch sums, Kdown/Kup/Kside/Ldown/Lside, Tmrt and UTCI
 * for hour ``hour`` (replaces Solweig_2022a_calc's per-pixel stages, model.py:141-251, with
 * gvf.cu, shortwave.cu, longwave.cu, flux.cu and utci.cu). full != 0 writes the FULL planes.
 * gvf_offsets_dev [18][S][2] (row, col) step shifts with S = second rounded up to a multiple of
 * 8, entries past ``second`` repeating the last step; gvf_bounds_dev [18][3] facesh bounds, as
 * prepare_gvf_geometry builds them (model.py:533-566). The neighbour record rec is [P] x 16 B. */
int sw_hour_pixels(const sw_domain *domains_dev, const int32_t *block_start_dev, int n_domains,
                   int total_blocks, int hour, const int32_t *gvf_offsets_dev,
                   const double *gvf_bounds_dev, int first_steps, int second_steps, int full,
                   void *stream);
/* Every hour in [hour_begin, hour_end): the two passes above back to back on one stream
 * (the day loop of series.py:445-469 without host round trips).                           */
int sw_hours_forward(const sw_domain *domains_dev, const int32_t *block_start_dev, int n_domains,
                     int total_blocks, int hour_begin, int hour_end, const int32_t *gvf_offsets_dev,
                     const double *gvf_bounds_dev, int first_steps, int second_steps, void *stream);
/* The same hours in two launches (replaces the same day loop, series.py:445-469): the shadow
 * pass of every hour, then one pixel pass in which each workgroup runs its tile through the hours
 * in order with the lag state in registers (read from and written back to the lag planes once).
 * Each domain's rec and masks planes must hold hour_end slots of P entries ([hours][P] x 8 B and
 * [hours][P] x 1 B); hour h uses slot h.                                                    */
int sw_day_forward(const sw_domain *domains_dev, const int32_t *block_start_dev, int n_domains,
                   int total_blocks, int hour_begin, int hour_end, const int32_t *gvf_offsets_dev,
                   const double *gvf_bounds_dev, int first_steps, int second_steps, void *stream);
/* The two launches of sw_day_forward on their own: the shadow pass of every hour of the range
 * (one launch, a workgroup row per hour), and the pixel pass of the range.                   */
int sw_day_shadow(const sw_domain *domains_dev, const int32_t *block_start_dev, int n_domains,
                  int total_blocks, int hour_begin, int hour_end, void *stream);
int sw_day_pixels(const sw_domain *domains_dev, const int32_t *block_start_dev, int n_domains,
                  int total_blocks, int hour_begin, int hour_end, const int32_t *gvf_offsets_dev,
                  const double *gvf_bounds_dev, int first_steps, int second_steps, void *stream);
/* sw_day_pixels with at most groups_per_cu (2..8) of its workgroups on one CU (the launch asks
 * for the LDS that enforces it), for a pixel pass that runs beside another day's sw_day_shadow
 * on a second stream: the shadow pass's waves then find registers on every CU (the same day
 * loop, series.py:445-469, with consecutive days overlapped). Results equal sw_day_pixels.   */
int sw_day_pixels_capped(const sw_domain *domains_dev, const int32_t *block_start_dev, int n_domains,
                         int total_blocks, int hour_begin, int hour_end, const int32_t *gvf_offsets_dev,
                         const double *gvf_bounds_dev, int first_steps, int second_steps,
                         int groups_per_cu, void *stream);
/* The diffuse patch sums of every daytime hour of a day, for every pixel of every domain whose
 * record has diffuse_tab and diffuse set (others are skipped): the per-patch luminance loop of
 * umep_solweig.py:148-154 / shortwave.cu:99-107 and the diffuse Kside term, shortwave.cu:155-157,
 * as one launch of int8 matrix products (visibility bits x table digits, exact integer sums, one
 * rounding to FP32 per sum). Run it after the day's table is in place and before the pixel
 * passes (sw_hours_forward, sw_day_forward, sw_day_pixels, sw_hour_pixels) that read the sums;
 * n_hours (1..65535) is the day's hour count the table was built for.                        */
int sw_day_diffuse(const sw_domain *domains_dev, const int32_t *block_start_dev, int n_domains,
                   int total_blocks, int n_hours, void *stream);

/* ----- adjoint of J = mean over ground pixels of UTCI at one objective hour -------------------
 * (no counterpart in the reference: the differentiable form of the per-step chain
 * model.py:141-251 + utci.cu, with shadow and sky-view masks held fixed). Per hour the device
 * returns dJ/d(hour-record value) for the values below, per-patch sums for the luminance and
 * diffuse tables, and accumulates dJ/d(albedo map) and dJ/d(emissivity map).                  */
enum {
    SWA_TA = 0, SWA_PSI, SWA_CITGG, SWA_LWALL, SWA_ALBEDO_B, SWA_RADI, SWA_RADD, SWA_RADG,
    SWA_FLUX4, SWA_SW_SUNLIT, SWA_SW_SHADED, SWA_SKY_A, SWA_SKY_B, SWA_LW_AMBIENT, SWA_LW_SUNLIT,
    SWA_TA32, SWA_UTCI0,                      /* 7 UTCI coefficients                                */
    SWA_TRANS_STATIC = SWA_UTCI0 + 7,         /* dJ/d transmissivity through svfbuveg           */
    SWA_SLOTS = 24,
    SWA_PARTIAL_SLOTS = SWA_SLOTS + 2         /* the hour's slots, then the objective sum and count */
};
/* per-patch sums at the objective hour: sum over pixels of lambda_Kdown (resp. lambda_Kside)
 * over the pixels that see patch k as sky (resp. through vegetation only)                   */
enum { SWA_PATCH_SKY_KDOWN = 0, SWA_PATCH_VEG_KDOWN, SWA_PATCH_SKY_KSIDE, SWA_PATCH_VEG_KSIDE,
       SWA_PATCH_SUMS = 4 };

/* Device buffers of the adjoint of one domain (all device pointers; 16 x 8 bytes). */
typedef struct sw_adjoint {
    float *lam_state;      /* [5][P] adjoint of the Lup lag states (all, E, S, W, N)        */
    float *lam_rec;        /* [3][P] adjoint of the neighbour-record Lup, albedo x shadow, albedo */
    float *coef;           /* [18][P][4] per-direction ground-view coefficients of the hour  */
    float *grad_alb;       /* [P] dJ/d albedo map, accumulated                                 */
    float *grad_emis;      /* [P] dJ/d emissivity map, accumulated                             */
    float *lam_patch;      /* [2][P] lambda_Kdown, lambda_Kside at the objective hour           */
    const float *svfveg;   /* [P] vegetation sky view factor (FP32)                           */
    double *scalars;       /* [n_hours][SWA_SLOTS]                                             */
    double *patch;         /* [SWA_PATCH_SUMS][153]                                            */
    double *objective;     /* [2] sum of UTCI over ground pixels at the objective hour, count */
    double *partial;       /* [SWA_PARTIAL_SLOTS][total_blocks] per-workgroup sums (work area;
                              NULL: one FP64 atomic per workgroup and slot instead, slower)      */
    int64_t reserved[5];
} sw_adjoint;

/* One hour of the backward sweep for a single domain: the shadow pass of ``hour`` must have
 * run (its neighbour records are read). At hour == objective_hour the forward of the hour is
 * recomputed (the lag states must hold the previous hour's) and the adjoint seeded with
 * 1/count per ground pixel (inv_count); earlier hours receive the lag-state adjoint only.
 * Requires first_steps == 1.                                                                 */
int sw_adjoint_hour(const sw_domain *domain_dev, const int32_t *block_start_dev, int total_blocks,
                    int hour, int objective_hour, const sw_adjoint *adjoint_dev,
                    const int32_t *gvf_offsets_dev, const double *gvf_bounds_dev, int first_steps,
                    int second_steps, double inv_count, void *stream);
/* Forward of hours [0, objective_hour) then the backward sweep objective_hour .. first_hour
 * (the first daytime hour, where the lag resets), recomputing each hour's shadow pass.
 * Buffers of adjoint_dev must be zeroed by the caller; lag states reset by the caller.      */
int sw_day_adjoint(const sw_domain *domain_dev, const int32_t *block_start_dev, int total_blocks,
                   int objective_hour, int first_hour, const sw_adjoint *adjoint_dev,
                   const int32_t *gvf_offsets_dev, const double *gvf_bounds_dev, int first_steps,
                   int second_steps, double inv_count, void *stream);

/* ----- per-step sub-models as standalone elementwise calls ------------------------------------
 * fields[k] is a device pointer to counts[k] values (1 = broadcast, or n); the working precision
 * is real_bytes (4: float, 8: double) for every field except the FP64 svfalfa of the wedge;
 * params are host FP64 values; out_dev receives n_out planes of n values (working precision).
 *   sw_cylindric_wedge     fields svfalfa (FP64);                 params zen;            1 plane
 *                          (model.cylindric_wedge, model.py:736-739; umep_solweig.py:383-426)
 *   sw_wave_delay          fields current, state;                 params first, update,
 *                          weight; planes delayed, next state  (model.TsWaveDelay_2015a, 742-756)
 *   sw_kup                 fields svfbuveg, F_sh, gvfalb (all, E, S, W, N), gvfalbnosh (all, E,
 *                          S, W, N); params radI, radD, radG, altitude, albedo_b; 5 planes Kup,
 *                          KupE, KupS, KupW, KupN       (model.Kup_veg_2015a, model.py:759-768)
 *   sw_lside               fields svf E/S/W/N, svfveg E/S/W/N, svfaveg E/S/W/N, Tw, Ldown, F_sh,
 *                          Lup E/S/W/N; params azimuth, altitude, Ta, SBC, ewall, esky, t, CI,
 *                          anisotropic; planes Least, Lsouth, Lwest, Lnorth
 *                                                       (model.Lside_veg_v2022a, model.py:771-781)
 *   sw_surface_temperature fields TgK, Tstart, TgK_wall, Tstart_wall, TmaxLST, TmaxLST_wall;
 *                          params altmax, frac(dectime), SNUP, CI_TgG, landcover; planes Tg,
 *                          Tgwall                   (model.surface_temperature, model.py:784-793)
 *   sw_assemble_tmrt       fields Kdown, Kup, Knorth, Keast, Ksouth, Kwest, Ldown, Lup, Lnorth,
 *                          Least, Lsouth, Lwest, KsideI, Kside, Lside; params absK, absL, Fside,
 *                          Fup, Fcyl, cyl, anisotropic, SBC; planes Tmrt, Sstr
 *                                                         (model.assemble_tmrt, model.py:796-805) */
int sw_cylindric_wedge(const void *const *fields, const long long *counts, int n_fields, long long n,
                       const double *params, int n_params, int real_bytes, void *out_dev, void *stream);
int sw_wave_delay(const void *const *fields, const long long *counts, int n_fields, long long n,
                  const double *params, int n_params, int real_bytes, void *out_dev, void *stream);
int sw_kup(const void *const *fields, const long long *counts, int n_fields, long long n,
           const double *params, int n_params, int real_bytes, void *out_dev, void *stream);
int sw_lside(const void *const *fields, const long long *counts, int n_fields, long long n,
             const double *params, int n_params, int real_bytes, void *out_dev, void *stream);
int sw_surface_temperature(const void *const *fields, const long long *counts, int n_fields, long long n,
                           const double *params, int n_params, int real_bytes, void *out_dev, void *stream);
int sw_assemble_tmrt(const void *const *fields, const long long *counts, int n_fields, long long n,
                     const double *params, int n_params, int real_bytes, void *out_dev, void *stream);
/* The elementwise terms the reference's step evaluates as CuPy expressions (replaces
 * model.py:120-126, 165-171, 183-195), one plane each, same conventions as above:
 *   op 0  shadow = sh - (1 - vegsh)(1 - psi)   fields sh, vegsh (FP64); params psi
 *   op 1  Kdown                                fields shadow, dRad, svfbuveg, F_sh;
 *                                              params radI, sin(altitude), albedo_b, radG, radD
 *   op 2  TgTemp = Tg shadow + Ta              fields Tg, shadow; params Ta
 *   op 3  night Lup                            fields emis, lc_grid (FP64);
 *                                              params Ta, SBC, landcover, water Lup
 *   op 4  NaN replaced by a value              fields x; params value
 *   op 5  a + b                                fields a, b; no params
 *   op 6  x rounded to the working precision   fields x (FP64); no params (the float32
 *                                              Tmrt that utci_calculator takes, utci.py:8-13)  */
int sw_step_terms(int op, const void *const *fields, const long long *counts, int n_fields, long long n,
                  const double *params, int n_params, int real_bytes, void *out_dev, void *stream);
/* dRad = radD sum_k diffsh[k][p] lum[k] (replaces model._diffuse, model.py:1019-1027, and the
 * diffuse kernel of shortwave.cu:98-107): diffsh_dev [patches][n] of the working precision,
 * lum_dev [patches] FP64.                                                                   */
int sw_diffuse(const void *diffsh_dev, const double *lum_dev, long long n, int patches, double radD,
               int real_bytes, void *out_dev, void *stream);

/* Static part of the ground view of a domain (the walk of gvf_2018a depends on the hour only
 * through Lup, albedo x shadow and the sunlit walls): per pixel and direction the number of
 * steps before the first non-ground sample, with bit 7 set where the pixel's wall faces away
 * from the direction (facesh == 0, umep_solweig.py:1554-1563) (len_dev [5][P] 32-bit words,
 * direction d in byte d % 4 of word d / 4, bytes 2 and 3 of word 4 zero), in the same layout the
 * number of steps whose shift stays inside the raster (edge_dev, may be NULL), and the
 * unshadowed-albedo sums of gvfalbnosh split as a + albedo_b * b over all 18 directions and the
 * E, S, W, N halves (nosh_dev [10][P]). ground_dev: [P] 0/1; alb_dev, walls_dev, aspect_dev
 * (degrees): [P]; offsets and bounds as sw_hour_pixels takes them.                          */
int sw_gvf_static(const uint8_t *ground_dev, const float *alb_dev, const float *walls_dev,
                  const float *aspect_dev, int rows, int cols, const int32_t *gvf_offsets_dev,
                  const double *gvf_bounds_dev, int first_steps, int second_steps, uint32_t *len_dev,
                  float *nosh_dev, uint32_t *edge_dev, void *stream);

/* The ground-view step shifts compiled into the library: out receives [18][n][2] (row, column)
 * for the first n <= 22 steps of the 18 directions (sunonsurface_2018a, umep_solweig.py:1477-1492).
 * With first_steps 1 and second_steps 22 (scale 1) sw_hour_pixels walks these shifts instead of
 * gvf_offsets_dev, which must then hold the same values (as prepare_gvf_geometry builds them).
 * No counterpart in the reference (it passes the table only).                                */
int sw_gvf_fixed_steps(int32_t *out, int n);

/* Ground view factors of one step for arbitrary caller fields (replaces model.gvf,
 * model.py:569-622, and its gvf_pack / gvf_compact / gvf kernels, gvf.cu:79-292; UMEP
 * gvf_2018a + sunonsurface_2018a, umep_solweig.py:583-663, 1400-1600). fields as above, 10 of
 * them: buildings, walls, aspect (radians), wallsun (FP64), shadow, Tg, Tgwall, emis, albedo
 * (working precision), lc_grid (FP64). params: Ta, ewall, SBC, albedo_b, landcover, Twater.
 * offsets_dev [18][step_stride][2] (row, col) and bounds_dev [18][3] as prepare_gvf_geometry
 * builds them (model.py:533-566). work_dev holds rows*cols*8 values of the working precision.
 * out_dev receives 18 planes: gvfLup, gvfalb, gvfalbnosh, the E, S, W, N triples, gvfSum,
 * gvfNorm, then Tg with water pixels (lc_grid == 3) replaced by Twater - Ta.            */
int sw_gvf(const void *const *fields, const long long *counts, int n_fields, int rows, int cols,
           const int32_t *offsets_dev, const double *bounds_dev, int step_stride, int first_steps,
           int second_steps, const double *params, int n_params, int real_bytes, void *work_dev,
           void *out_dev, void *stream);

/* sw_gvf without the work buffer: no counterpart in the reference. For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The per-pixel samples go to a library-owned static device array of
 * SW_GVF_SCRATCH_BYTES (64 MiB, held by every GPU the library runs on until the process ends):
 * rows * cols * 8 working-precision values must fit, i.e. at most 2097152 pixels in FP32 and
 * 1048576 in FP64; larger rasters return SW_ERR_LIMIT before any device work. Arguments and
 * results are those of sw_gvf. Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other host threads) race, so serialise them. The device
 * must be current (sw_set_device) on the calling thread.                                  */
#define SW_GVF_SCRATCH_BYTES (64ull << 20)
int sw_gvf_noalloc(const void *const *fields, const long long *counts, int n_fields, int rows, int cols,
                   const int32_t *offsets_dev, const double *bounds_dev, int step_stride, int first_steps,
                   int second_steps, const double *params, int n_params, int real_bytes, void *out_dev,
                   void *stream);

/* Visibility flags of a caller's patch geometry (replaces prepare_geometry's pack,
 * model.py:273-306, geometry.cu:1-8): shmat, vegshmat, vbshvegshmat are [n][patches] device
 * arrays of mask_bytes-sized values (1: uint8, 4: float, 8: double). flags_dev receives
 * [4][ceil(patches / 32)][n] words; bit k % 32 of word [f][k / 32][p] is flag f of patch k at
 * pixel p: 0 sky (sh == 1 and veg == 1), 1 vegetation (veg == 0 or vb == 0), 2 wall
 * ((1 - sh) * vb == 1), 3 obstructed (any of the three == 0).                             */
int sw_pack_visibility(const void *shmat_dev, const void *vegshmat_dev, const void *vbshvegshmat_dev,
                       int mask_bytes, long long n, int patches, uint32_t *flags_dev, void *stream);
/* Anisotropic-sky longwave of one step (replaces model.longwave, model.py:367-392, and
 * longwave.cu; UMEP Lcyl_v2022a / define_patch_characteristics, umep_solweig.py:946-1359).
 * table_dev: [patches][20] FP64 per-patch coefficients built by the host for the call (columns:
 * sky Ldown, sky Lside, vegetation x cos, x sin, shaded wall x cos, x sin, steradian, cos alt,
 * sin alt, direction cosines E, W, N, S, yi, sunlit-below and shaded-from thresholds of
 * tan(asvf) + yi, sun-test eligibility). fields: Lup, Tgwall (working precision), tan(asvf)
 * (FP64); params: Ta, ewall, SBC. out_dev: 6 planes Ldown, Lside, Least, Lwest, Lnorth, Lsouth. */
int sw_longwave(const uint32_t *flags_dev, int patches, long long n, const double *table_dev,
                const void *const *fields, const long long *counts, int n_fields, const double *params,
                int n_params, int real_bytes, void *out_dev, void *stream);
/* Shortwave on the body's sides (replaces model.Kside_veg_v2022a, model.py:443-506, and
 * shortwave.cu; UMEP umep_solweig.py:666-911), every (cyl, anisotropic_diffuse) branch.
 * diffsh_dev: [patches][n] working-precision diffuse transmissivities (prepare_shortwave's
 * transpose, model.py:401-411); table_dev: [patches][20] FP64 (lumChi, cylinder incidence,
 * steradian, box incidences E, S, W, N, sunlit-surface and shaded-surface terms, reflection
 * cosines E, S, W, N, yi, the two thresholds, sector flags). flags, diffsh and table may be NULL
 * when anisotropic_diffuse is 0. fields: shadow, F_sh, KupE, KupS, KupW, KupN, svfE, svfS, svfW,
 * svfN, svfEveg, svfSveg, svfWveg, svfNveg (working precision), tan(asvf) (FP64). params (16):
 * radI, radD, radG, cos(altitude), psi, albedo, sin of the sun's azimuth on faces E, S, W, N,
 * those faces lit (0/1), cyl, anisotropic_diffuse. out_dev: 7 planes Keast, Ksouth, Kwest,
 * Knorth, KsideI, KsideD, Kside.                                                          */
int sw_kside(const uint32_t *flags_dev, const void *diffsh_dev, int patches, long long n,
             const double *table_dev, const void *const *fields, const long long *counts, int n_fields,
             const double *params, int n_params, int real_bytes, void *out_dev, void *stream);

/* Wall heights and wall aspect from a DSM (replaces walls.calculate_dsm_walls, walls.py:24-142,
 * and its wall_heights / wall_directions kernels, geometry.cu:15-63). dsm_dev: FP64
 * [rows][cols]; threshold: minimum wall height; step: np.gradient spacing (1 / (1 / pixel
 * size)). Templates (host-built, walls.py:33-56): n_angles (180, or 0 when the raster is not
 * larger than 2 radius + 2 on both sides) rotated line templates, line_dev holding (row, col)
 * offset pairs of angle k at [line_start[k], line_start[k+1]); side_dev the two-sided
 * templates, angle k's halves at [side_start[3k], side_start[3k+1]) and
 * [side_start[3k+1], side_start[3k+2]). Offsets stay within +-radius. occupancy_dev: rows*cols
 * bytes of scratch. Outputs FP64 walls_dev (m) and aspect_dev (degrees clockwise from north).  */
int sw_dsm_walls(const double *dsm_dev, int rows, int cols, double threshold, double step, int radius,
                 int n_angles, const int32_t *line_dev, const int32_t *line_start_dev, const int32_t *side_dev,
                 const int32_t *side_start_dev, uint8_t *occupancy_dev, double *walls_dev, double *aspect_dev,
                 void *stream);

/* Solar position of forcing rows in FP64 (replaces Solweig_2015a_metdata_noload and
 * sun_position, umep_solweig.py:2412-2503, 2529-3521: the NREL SPA with delta_t = 0 and fixed
 * refraction). met_dev [n_rows][met_cols] holds the forcing rows (year, day of year, hour,
 * minute, ...; met_cols >= 4). The rows are taken in independent chunks of chunk_rows (0: one
 * chunk), each as one call of the reference: a row's sun is sampled half the chunk's first time
 * step before its stamp, a zenith in (89, 90] becomes 89, and the day's maximum altitude is
 * searched at the chunk's first row and at every midnight row. out_dev [6][n_rows]: altitude
 * (deg), azimuth (deg), zenith (rad), day of year, dectime, altmax (deg).                   */
int sw_sun_position(const double *met_dev, int n_rows, int met_cols, int chunk_rows, double latitude,
                    double longitude, double altitude_m, double utc, double *out_dev, void *stream);

/* UTCI of a Tmrt plane for scalar weather (replaces utci.utci_calculator's scalar-weather
 * path, utci.py:22-33, utci.cu:243-523): ``coefficients`` are 7 host FP64 values c_k with
 * UTCI = sum_k c_k (Tmrt - ta32)^k; sentinel != 0 writes -999; ground_dev (nullable) NaNs
 * pixels where it is 0 (series.py:468-469).                                                */
int sw_utci_scalar(const float *tmrt_dev, long long n, const double *coefficients, double ta32,
                   int sentinel, const uint8_t *ground_dev, float *out_dev, void *stream);

/* Streaming device copy (float4 loads/stores, bytes and addresses 16-byte aligned): the plain
 * kernel of known traffic used to calibrate HBM byte counters. No reference counterpart.       */
int sw_copy(void *dst_dev, const void *src_dev, size_t bytes, void *stream);

/* hipGraph capture of a launch sequence on ``stream`` (the captured day loop).            */
int sw_graph_begin(void *stream);
int sw_graph_end(void *stream, void **graph_exec);
int sw_graph_launch(void *graph_exec, void *stream);
int sw_graph_destroy(void *graph_exec);

#ifdef __cplusplus
}
#endif
#endif
