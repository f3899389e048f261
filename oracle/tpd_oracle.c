/*
 * CPU oracle for the Gaussian rasterization path of ndming/torpedo.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, smoke() and bench.py's
 * cpu_baseline leg load it to produce expected outputs.  The product library
 * (torpedo_amd/csrc) never links, loads or calls it.
 *
 * What it restates (reference files under torpedo/):
 *   camera      rendering/src/Camera.cpp:3-16, extension/src/PerspectiveCamera.cpp:3-23,
 *               extension/include/torpedo/extension/PerspectiveCamera.h:26-31,
 *               math/include/torpedo/math/common.h:11-40, vec3.h:230-262, vec4.h:248-276,
 *               mat4.h:256-264, volumetric/src/GaussianEngine.cpp:764-775
 *   preprocess  volumetric/assets/gaussian/project.slang:33-90, splat/common.slang:35-119,
 *               splat/volume.slang:3-63
 *   scan        prefix.slang:38-159 (exclusive scan of tile counts, total = tilesRendered)
 *   duplicate   keygen.slang:21-54
 *   sort        radix-*.slang + GaussianEngine.cpp:819-841 (stable LSD sort of (key, value))
 *   ranges      range.slang:16-35 + the fillBuffer clear at GaussianEngine.cpp:843-848
 *   blend       blend.slang:22-105
 *
 * exp(): the blend needs exp() in FP32.  The reference leaves it to the Vulkan driver, which
 * cannot run here; this oracle and the HIP kernels both use tpo_expf below (Cody-Waite
 * argument reduction, W. J. Cody and W. Waite, "Software Manual for the Elementary Functions",
 * 1980, followed by a degree-5 polynomial of e^r fitted for relative error and evaluated with
 * fused multiply-adds), so that threshold decisions (alpha < 1/255, T < 1e-4) cannot flip between
 * CPU and GPU.
 *
 * Parity status of each function is listed in the repository's design notes; the camera
 * functions are additionally pinned bit-for-bit against the reference's own math headers
 * compiled by oracle/ref_camera.sh.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define CAT(a, b) a##b
#define XCAT(a, b) CAT(a, b)

/* Saturating truncation to int32, NaN -> 0 (the GPU conversion's behaviour, made explicit). */
static inline int32_t sat_i32(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (int32_t)(-2147483647 - 1);
    return (int32_t)v;
}

/* exp(x) in FP32 for x <= 0 (the blend's domain).  Inputs below -86 (and NaN) return 0: every
 * caller compares opacity * exp(x) against 1/255, far above e^-86. */
float tpo_expf(float x) {
    if (!(x >= -86.0f)) return 0.0f;
    if (x > 88.0f) return INFINITY;
    /* n = round(x / ln 2) from the low bits of x log2(e) + 1.5 * 2^23, r = x - n ln 2 (Cody-Waite,
     * exact), e^r by a degree-5 relative-error fit on [-ln2/2, ln2/2], 2^n added to the exponent */
    union { float f; uint32_t u; } t, p;
    t.f = fmaf(x, 1.44269504088896341f, 12582912.0f);
    const float n = t.f - 12582912.0f;
    float r = fmaf(n, -0.693145751953125f, x);
    r = fmaf(n, -1.42860682030941723212e-6f, r);
    p.f = 0.008286180905997753f;
    p.f = fmaf(p.f, r, 0.04189886897802353f);
    p.f = fmaf(p.f, r, 0.16667655110359192f);
    p.f = fmaf(p.f, r, 0.49999144673347473f);
    p.f = fmaf(p.f, r, 0.9999997615814209f);
    p.f = fmaf(p.f, r, 1.0f);
    p.u += t.u << 23;
    return p.f;
}

void tpo_expf_vec(const float* x, float* y, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) y[i] = tpo_expf(x[i]);
}

/* ---------------------------------------------------------------- threads */

typedef void (*range_fn)(void* arg, uint64_t lo, uint64_t hi);
typedef struct { range_fn fn; void* arg; uint64_t lo, hi; } chunk_t;

static void* chunk_main(void* p) {
    chunk_t* c = (chunk_t*)p;
    c->fn(c->arg, c->lo, c->hi);
    return NULL;
}

static void parallel_for(range_fn fn, void* arg, uint64_t n, int nthreads) {
    if (nthreads <= 1 || n < 1024) { fn(arg, 0, n); return; }
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    chunk_t ch[256];
    for (int t = 0; t < nthreads; ++t) {
        ch[t].fn = fn; ch[t].arg = arg;
        ch[t].lo = n * (uint64_t)t / (uint64_t)nthreads;
        ch[t].hi = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, chunk_main, &ch[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* Every float x with lo <= x <= hi (lo < hi <= 0, lo > -87): the largest relative error of tpo_expf
 * against the double-precision exp, and the number of neighbouring floats whose results decrease
 * (*down), over nthreads threads. */
typedef struct { uint32_t u0; uint64_t lo, hi; double max_rel; uint64_t down; } expsweep_t;
static void* expsweep_main(void* p) {
    expsweep_t* s = (expsweep_t*)p;
    double m = 0.0;
    uint64_t down = 0;
    float prev = 0.0f;
    for (uint64_t k = s->lo; k < s->hi; ++k) {
        union { uint32_t u; float f; } x;
        x.u = s->u0 - (uint32_t)k;  /* from lo upwards: negative floats grow as their bit pattern shrinks */
        const float y = tpo_expf(x.f);
        const double ref = exp((double)x.f);
        const double e = fabs((double)y - ref) / ref;
        if (e > m) m = e;
        if (k > s->lo && y < prev) ++down;
        prev = y;
    }
    s->max_rel = m;
    s->down = down;
    return NULL;
}

void tpo_expf_sweep(float lo, float hi, int nthreads, double* max_rel, uint64_t* down) {
    union { float f; uint32_t u; } a, b;
    a.f = lo;
    b.f = hi;
    if (b.u == 0u) b.u = 0x80000000u;  /* +0 -> -0 */
    const uint64_t n = (uint64_t)(a.u - b.u) + 1;  /* floats from lo up to hi */
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    expsweep_t ch[256];
    for (int t = 0; t < nthreads; ++t) {
        ch[t].u0 = a.u;
        ch[t].lo = n * (uint64_t)t / (uint64_t)nthreads;
        ch[t].hi = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        /* one float of overlap so that the monotonicity check also covers the chunk boundaries */
        if (t > 0) ch[t].lo -= 1;
        pthread_create(&th[t], NULL, expsweep_main, &ch[t]);
    }
    *max_rel = 0.0;
    *down = 0;
    for (int t = 0; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        if (ch[t].max_rel > *max_rel) *max_rel = ch[t].max_rel;
        *down += ch[t].down;
    }
}

/* ---------------------------------------------------------------- camera (host math) */

/* math/common.h:11-17: a*b - c*d with Kahan's FMA error term. */
static float dop(float a, float b, float c, float d) {
    const float cd = c * d;
    const float err = fmaf(-c, d, cd);
    const float r = fmaf(a, b, -cd);
    return r + err;
}
/* common.h:23-29 TwoSum */
static void two_sum(float a, float b, float* s, float* e) {
    const float sum = a + b;
    const float z = sum - a;
    *s = sum;
    *e = a - (sum - z) + (b - z);
}
/* common.h:35-40 TwoProduct */
static void two_prod(float a, float b, float* p, float* e) {
    const float m = a * b;
    *p = m;
    *e = fmaf(a, b, -m);
}
/* vec3.h:239-252 / vec4.h:248-266: compensated dot (Graillat, "Algorithm 6"). */
static float dot_comp(const float* a, const float* b, int n) {
    float xx, e0, yy, e1, s, e3;
    two_prod(a[0], b[0], &xx, &e0);
    two_prod(a[1], b[1], &yy, &e1);
    two_sum(xx, yy, &s, &e3);
    float err = e0 + (e3 + e1);
    for (int k = 2; k < n; ++k) {
        float pk, ek, sk, es;
        two_prod(a[k], b[k], &pk, &ek);
        two_sum(s, pk, &sk, &es);
        if (k == n - 1) return sk + (err + (es + ek));
        err = err + (es + ek);
        s = sk;
    }
    return s + err;
}
static void cross_comp(const float* a, const float* b, float* o) {
    o[0] = dop(a[1], b[2], a[2], b[1]);
    o[1] = dop(a[2], b[0], a[0], b[2]);
    o[2] = dop(a[0], b[1], a[1], b[0]);
}
/* vec3.h:254-262: v * (1 / sqrt(dot(v, v))) (vec3 operator/ multiplies by the reciprocal). */
static void normalize3(const float* v, float* o) {
    const float inv = 1.0f / sqrtf(dot_comp(v, v, 3));
    o[0] = v[0] * inv; o[1] = v[1] * inv; o[2] = v[2] * inv;
}

/* GaussianPoint::fromModel's conversion of one PLY vertex (volumetric/src/GaussianGeometry.cpp:108-119):
 * sigmoid opacity, (rot_1, rot_2, rot_3, rot_0) normalised by v * (1 / sqrt(compensated dot))
 * (math/vec4.h:248-275), exp scales with modifier 1, nfeat SH features copied (rest zero). */
void tpo_gaussian_from_ply(const float* xyz, float opacity_logit, const float* rot_wxyz, const float* log_scale,
                           const float* features, uint32_t nfeat, float* out) {
    memset(out, 0, 60 * sizeof(float));
    out[0] = xyz[0]; out[1] = xyz[1]; out[2] = xyz[2];
    out[3] = 1.0f / (1.0f + expf(-opacity_logit));
    const float q[4] = {rot_wxyz[1], rot_wxyz[2], rot_wxyz[3], rot_wxyz[0]};
    const float inv = 1.0f / sqrtf(dot_comp(q, q, 4));
    for (int i = 0; i < 4; ++i) out[4 + i] = q[i] * inv;
    for (int i = 0; i < 3; ++i) out[8 + i] = expf(log_scale[i]);
    out[11] = 1.0f;
    for (uint32_t k = 0; k < nfeat && k < 48; ++k) out[12 + k] = features[k];
}

void tpo_look_at(const float* eye, const float* center, const float* up, float* view) {
    float d[3] = {center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]};
    float z[3], x[3], y[3], t[3];
    normalize3(d, z);
    cross_comp(z, up, t); normalize3(t, x);
    cross_comp(z, x, t); normalize3(t, y);
    const float r[16] = {
        x[0], x[1], x[2], -dot_comp(x, eye, 3),
        y[0], y[1], y[2], -dot_comp(y, eye, 3),
        z[0], z[1], z[2], -dot_comp(z, eye, 3),
        0.f, 0.f, 0.f, 1.f,
    };
    memcpy(view, r, sizeof(r));
}

/* PerspectiveCamera.cpp:9-23 (reversed z). */
void tpo_perspective(uint32_t w, uint32_t h, float fy, float znear, float zfar, float* proj) {
    const float aspect = (float)w / (float)h;
    const float fx = fy / aspect;
    const float za = znear / (znear - zfar);
    const float zb = znear * zfar / (zfar - znear);
    const float r[16] = {fx, 0.f, 0.f, 0.f, 0.f, fy, 0.f, 0.f, 0.f, 0.f, za, zb, 0.f, 0.f, 1.f, 0.f};
    memcpy(proj, r, sizeof(r));
}

/* PerspectiveCamera.cpp:3-7 */
float tpo_fy_from_fov(float degrees) {
    const float fov = degrees * 3.14159265358979323846f / 180.f;
    return 1.f / tanf(fov * 0.5f);
}

/* mat4.h:256-264 compensated product. */
void tpo_mat4_mul(const float* a, const float* b, float* out) {
    float r[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const float col[4] = {b[0 * 4 + j], b[1 * 4 + j], b[2 * 4 + j], b[3 * 4 + j]};
            r[i * 4 + j] = dot_comp(a + i * 4, col, 4);
        }
    memcpy(out, r, sizeof(r));
}

/* GaussianEngine.cpp:764-775 updateCameraBuffer: projView = P * V, focal = (P[0][0], P[1][1]). */
void tpo_camera_upload(const float* view, const float* proj, float* proj_view, float* focal_ndc) {
    tpo_mat4_mul(proj, view, proj_view);
    focal_ndc[0] = proj[0];
    focal_ndc[1] = proj[5];
}

/* ---------------------------------------------------------------- precision-generic stages */

#define REAL float
#define SUFFIX _f32
#define K(x) ((float)(x))
#define SQRT sqrtf
#define CEIL ceilf
#define RINT rintf
#define EXP tpo_expf
#include "oracle_core.inc"
#undef REAL
#undef SUFFIX
#undef K
#undef SQRT
#undef CEIL
#undef RINT
#undef EXP

#define REAL double
#define SUFFIX _f64
#define K(x) ((double)(x))
#define SQRT sqrt
#define CEIL ceil
#define RINT rint
#define EXP exp
#include "oracle_core.inc"
#undef REAL
#undef SUFFIX
#undef K
#undef SQRT
#undef CEIL
#undef RINT
#undef EXP

/* ---------------------------------------------------------------- exported stages (FP32) */

static void entity_mats_f32(const float* view, const float* proj_view, const float* models, uint32_t n_ent,
                            float* VMs, float* PVMs) {
    static const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    for (uint32_t e = 0; e < (n_ent ? n_ent : 1u); ++e) {
        const float* M = models ? models + 16 * e : I;
        mat4_mul_plain_f32(view, M, VMs + 16 * e);
        mat4_mul_plain_f32(proj_view, M, PVMs + 16 * e);
    }
}

/*
 * Preprocess N Gaussians (60 floats each).  ent_idx/models may be NULL (one identity entity).
 * Outputs (all length-N arrays, culled Gaussians are all-zero): color[3N], texel[4N], copac[4N],
 * tiles[N], rect[4N].
 */
void tpo_preprocess_f32(const float* gauss, uint32_t n, const uint32_t* ent_idx, const float* models, uint32_t n_ent,
                        const float* view, const float* proj_view, const float* focal_ndc, uint32_t W, uint32_t H,
                        uint32_t sh_degree, float* color, float* texel, float* copac, uint32_t* tiles, int32_t* rect,
                        int nthreads) {
    const uint32_t ne = n_ent ? n_ent : 1u;
    float* VMs = (float*)malloc(sizeof(float) * 16 * ne);
    float* PVMs = (float*)malloc(sizeof(float) * 16 * ne);
    entity_mats_f32(view, proj_view, models, n_ent, VMs, PVMs);
    splat_t_f32* out = (splat_t_f32*)malloc(sizeof(splat_t_f32) * (n ? n : 1));
    project_job_t_f32 job = {gauss, n, models ? ent_idx : NULL, VMs, PVMs, view,
                             (0.5f * (float)W) * focal_ndc[0], (0.5f * (float)H) * focal_ndc[1], W, H,
                             sh_degree > 3 ? 3 : sh_degree, out};
    parallel_for(project_range_f32, &job, n, nthreads);
    for (uint32_t i = 0; i < n; ++i) {
        memcpy(color + 3 * (size_t)i, out[i].color, 12);
        memcpy(texel + 4 * (size_t)i, out[i].texel, 16);
        memcpy(copac + 4 * (size_t)i, out[i].copac, 16);
        tiles[i] = out[i].tiles;
        if (rect) memcpy(rect + 4 * (size_t)i, out[i].rect, 16);
    }
    free(out); free(VMs); free(PVMs);
}

/* prefix.slang:38-159: exclusive scan; returns the total (tilesRendered). */
uint64_t tpo_scan(const uint32_t* tiles, uint32_t n, uint32_t* offsets) {
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        offsets[i] = (uint32_t)acc;
        acc += tiles[i];
    }
    return acc;
}

typedef struct {
    const float* texel; const uint32_t* offsets; uint32_t gx, gy;
    uint64_t* keys; uint32_t* vals;
} dup_job_t;

static void dup_range(void* arg, uint64_t lo, uint64_t hi) {
    dup_job_t* j = (dup_job_t*)arg;
    for (uint64_t i = lo; i < hi; ++i) {
        const float* t = j->texel + 4 * i;
        if (t[3] <= 0.0f) continue;
        int32_t rect[4];
        bounding_rect_f32(t[0], t[1], t[3], j->gx, j->gy, rect);
        uint32_t depth_bits;
        memcpy(&depth_bits, &t[2], 4);
        uint64_t off = j->offsets[i];
        for (int32_t y = rect[1]; y < rect[3]; ++y)
            for (int32_t x = rect[0]; x < rect[2]; ++x) {
                j->keys[off] = ((uint64_t)((uint32_t)y * j->gx + (uint32_t)x) << 32) | depth_bits;
                j->vals[off] = (uint32_t)i;
                ++off;
            }
    }
}

/* keygen.slang:21-54 */
void tpo_duplicate(const float* texel, const uint32_t* offsets, uint32_t n, uint32_t W, uint32_t H, uint64_t* keys,
                   uint32_t* vals, int nthreads) {
    dup_job_t job = {texel, offsets, (W + 15u) / 16u, (H + 15u) / 16u, keys, vals};
    parallel_for(dup_range, &job, n, nthreads);
}

/* Stable LSD radix sort of (key, value) by key, 8-bit digits, passes whose digit is constant
 * over all keys are skipped.  Equivalent to the reference's 2-bit LSD passes over the low
 * 2*ceil((bitlen(tiles)+32)/2) bits, since all higher key bits are zero. */
typedef struct {
    const uint64_t* kin; const uint32_t* vin; uint64_t* kout; uint32_t* vout;
    uint64_t n; int shift, nthreads; uint64_t (*hist)[256];
} rs_job_t;

static void rs_hist(void* arg, uint64_t lo, uint64_t hi) {
    rs_job_t* j = (rs_job_t*)arg;
    const uint64_t t = lo;
    const uint64_t a = j->n * t / (uint64_t)j->nthreads, b = j->n * (t + 1) / (uint64_t)j->nthreads;
    (void)hi;
    uint64_t* h = j->hist[t];
    memset(h, 0, sizeof(uint64_t) * 256);
    for (uint64_t i = a; i < b; ++i) h[(j->kin[i] >> j->shift) & 255u]++;
}
static void rs_scatter(void* arg, uint64_t lo, uint64_t hi) {
    rs_job_t* j = (rs_job_t*)arg;
    const uint64_t t = lo;
    const uint64_t a = j->n * t / (uint64_t)j->nthreads, b = j->n * (t + 1) / (uint64_t)j->nthreads;
    (void)hi;
    uint64_t* h = j->hist[t];
    for (uint64_t i = a; i < b; ++i) {
        const uint64_t o = h[(j->kin[i] >> j->shift) & 255u]++;
        j->kout[o] = j->kin[i];
        j->vout[o] = j->vin[i];
    }
}
static void run_per_thread(range_fn fn, void* arg, int nthreads) {
    if (nthreads <= 1) { fn(arg, 0, 1); return; }
    pthread_t th[256];
    chunk_t ch[256];
    for (int t = 0; t < nthreads; ++t) {
        ch[t].fn = fn; ch[t].arg = arg; ch[t].lo = (uint64_t)t; ch[t].hi = (uint64_t)t + 1;
        pthread_create(&th[t], NULL, chunk_main, &ch[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

void tpo_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t n, int nthreads) {
    if (n < 2) return;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (n < 65536) nthreads = 1;
    uint64_t* kt = (uint64_t*)malloc(sizeof(uint64_t) * n);
    uint32_t* vt = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint64_t (*hist)[256] = malloc(sizeof(uint64_t) * 256 * (size_t)nthreads);
    uint64_t* kin = keys; uint32_t* vin = vals; uint64_t* kout = kt; uint32_t* vout = vt;
    for (int shift = 0; shift < 64; shift += 8) {
        rs_job_t j = {kin, vin, kout, vout, n, shift, nthreads, hist};
        run_per_thread(rs_hist, &j, nthreads);
        uint64_t tot[256] = {0};
        for (int t = 0; t < nthreads; ++t)
            for (int d = 0; d < 256; ++d) tot[d] += hist[t][d];
        int trivial = 0;
        for (int d = 0; d < 256; ++d) if (tot[d] == n) trivial = 1;
        if (trivial) continue;
        uint64_t acc = 0;
        for (int d = 0; d < 256; ++d)
            for (int t = 0; t < nthreads; ++t) {
                const uint64_t c = hist[t][d];
                hist[t][d] = acc;
                acc += c;
            }
        run_per_thread(rs_scatter, &j, nthreads);
        uint64_t* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    if (kin != keys) {
        memcpy(keys, kin, sizeof(uint64_t) * n);
        memcpy(vals, vin, sizeof(uint32_t) * n);
    }
    free(kt); free(vt); free(hist);
}

/* range.slang:16-35 after the zero fill: ranges[2*tile] = start, ranges[2*tile+1] = end. */
void tpo_ranges(const uint64_t* keys, uint64_t n, uint32_t ntiles, uint32_t* ranges) {
    memset(ranges, 0, sizeof(uint32_t) * 2 * (size_t)ntiles);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t cur = (uint32_t)(keys[i] >> 32);
        if (i == 0) {
            ranges[2 * (size_t)cur] = 0;
        } else {
            const uint32_t prev = (uint32_t)(keys[i - 1] >> 32);
            if (cur != prev) {
                ranges[2 * (size_t)prev + 1] = (uint32_t)i;
                ranges[2 * (size_t)cur] = (uint32_t)i;
            }
        }
        if (i == n - 1) ranges[2 * (size_t)cur + 1] = (uint32_t)(i + 1);
    }
}

typedef struct {
    const splat_t_f32* sp; const uint32_t* vals; const uint32_t* ranges;
    uint32_t W, H, gx; float* image; uint8_t* rgba8;
} blend_job_t;

static void blend_range(void* arg, uint64_t lo, uint64_t hi) {
    blend_job_t* j = (blend_job_t*)arg;
    for (uint64_t t = lo; t < hi; ++t) {
        const uint32_t s = j->ranges[2 * t], e = j->ranges[2 * t + 1];
        blend_tile_f32(j->sp, j->vals + s, e - s, (uint32_t)(t % j->gx), (uint32_t)(t / j->gx), j->W, j->H, j->image,
                       j->rgba8);
    }
}

/* blend.slang:22-105: image is W*H*3 floats, rgba8 is W*H*4 bytes (either may be NULL). */
void tpo_blend_f32(const float* color, const float* texel, const float* copac, uint32_t n, const uint32_t* vals,
                   const uint32_t* ranges, uint32_t W, uint32_t H, float* image, uint8_t* rgba8, int nthreads) {
    splat_t_f32* sp = (splat_t_f32*)calloc(n ? n : 1, sizeof(splat_t_f32));
    for (uint32_t i = 0; i < n; ++i) {
        memcpy(sp[i].color, color + 3 * (size_t)i, 12);
        memcpy(sp[i].texel, texel + 4 * (size_t)i, 16);
        memcpy(sp[i].copac, copac + 4 * (size_t)i, 16);
    }
    const uint32_t gx = (W + 15u) / 16u, gy = (H + 15u) / 16u;
    blend_job_t job = {sp, vals, ranges, W, H, gx, image, rgba8};
    parallel_for(blend_range, &job, (uint64_t)gx * gy, nthreads);
    free(sp);
}

/*
 * The FP32 blend's discrete decisions (blend.slang:54-101): for every pixel (row-major), the
 * sorted-list positions k of the splats it actually blended, in blend order.  counts[p] receives
 * the number per pixel; positions (capacity entries, may be NULL for a counting call) the
 * concatenated lists.  Returns the total.  Used to evaluate a dense FP64 reference on exactly the
 * (pixel, splat) pairs the FP32 frame blended.
 */
uint64_t tpo_blend_decisions_f32(const float* color, const float* texel, const float* copac, uint32_t n,
                                 const uint32_t* vals, const uint32_t* ranges, uint32_t W, uint32_t H,
                                 uint32_t* counts, uint32_t* positions, uint64_t capacity) {
    (void)color;
    (void)n;
    const uint32_t gx = (W + 15u) / 16u;
    uint64_t total = 0;
    for (uint32_t py = 0; py < H; ++py) {
        for (uint32_t px = 0; px < W; ++px) {
            const uint32_t t = (py / 16u) * gx + (px / 16u);
            float Tr = 1.0f;
            uint32_t cnt = 0;
            for (uint32_t k = ranges[2 * t]; k < ranges[2 * t + 1]; ++k) {
                const uint32_t g = vals[k];
                const float* tx = texel + 4 * (size_t)g;
                const float* cp = copac + 4 * (size_t)g;
                const float dx = tx[0] - (float)px, dy = tx[1] - (float)py;
                const float power = -0.5f * ((cp[0] * dx) * dx + (cp[2] * dy) * dy) - (cp[1] * dx) * dy;
                if (power > 0.0f) continue;
                float alpha = cp[3] * tpo_expf(power);
                alpha = alpha < 0.99f ? alpha : 0.99f;
                if (alpha < 1.0f / 255.0f) continue;
                if (Tr * (1.0f - alpha) < 0.0001f) break;
                Tr = Tr * (1.0f - alpha);
                if (positions && total < capacity) positions[total] = k;
                ++total;
                ++cnt;
            }
            counts[(size_t)py * W + px] = cnt;
        }
    }
    return total;
}

/* ---------------------------------------------------------------- whole pipelines */

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/*
 * The FP32 forward frame end to end (the CPU baseline).  Returns the pair count P.
 * stage_s[0..4] receive seconds spent in preprocess, scan+duplicate, sort, ranges, blend.
 */
uint64_t tpo_render_f32(const float* gauss, uint32_t n, const uint32_t* ent_idx, const float* models, uint32_t n_ent,
                        const float* view, const float* proj_view, const float* focal_ndc, uint32_t W, uint32_t H,
                        uint32_t sh_degree, float* image, uint8_t* rgba8, double* stage_s, int nthreads) {
    double t0 = now_s();
    float* color = (float*)malloc(sizeof(float) * 3 * (size_t)(n ? n : 1));
    float* texel = (float*)malloc(sizeof(float) * 4 * (size_t)(n ? n : 1));
    float* copac = (float*)malloc(sizeof(float) * 4 * (size_t)(n ? n : 1));
    uint32_t* tiles = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    uint32_t* offs = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    tpo_preprocess_f32(gauss, n, ent_idx, models, n_ent, view, proj_view, focal_ndc, W, H, sh_degree, color, texel,
                       copac, tiles, NULL, nthreads);
    double t1 = now_s();
    const uint64_t P = tpo_scan(tiles, n, offs);
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(P ? P : 1));
    uint32_t* vals = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(P ? P : 1));
    tpo_duplicate(texel, offs, n, W, H, keys, vals, nthreads);
    double t2 = now_s();
    tpo_sort_pairs(keys, vals, P, nthreads);
    double t3 = now_s();
    const uint32_t ntiles = ((W + 15u) / 16u) * ((H + 15u) / 16u);
    uint32_t* ranges = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (size_t)ntiles);
    tpo_ranges(keys, P, ntiles, ranges);
    double t4 = now_s();
    tpo_blend_f32(color, texel, copac, n, vals, ranges, W, H, image, rgba8, nthreads);
    double t5 = now_s();
    if (stage_s) {
        stage_s[0] = t1 - t0; stage_s[1] = t2 - t1; stage_s[2] = t3 - t2; stage_s[3] = t4 - t3; stage_s[4] = t5 - t4;
    }
    free(color); free(texel); free(copac); free(tiles); free(offs); free(keys); free(vals); free(ranges);
    return P;
}

/* FP64 twin: pairs are ordered by (tile, depth, index) with depth kept in double. */
typedef struct { uint32_t tile; double depth; uint32_t idx; } pair64_t;
static int cmp_pair64(const void* a, const void* b) {
    const pair64_t* x = (const pair64_t*)a;
    const pair64_t* y = (const pair64_t*)b;
    if (x->tile != y->tile) return x->tile < y->tile ? -1 : 1;
    if (x->depth != y->depth) return x->depth < y->depth ? -1 : 1;
    if (x->idx != y->idx) return x->idx < y->idx ? -1 : 1;
    return 0;
}

typedef struct {
    const splat_t_f64* sp; const uint32_t* order; const uint64_t* tstart; uint32_t W, H, gx; double* image;
} blend64_job_t;
static void blend64_range(void* arg, uint64_t lo, uint64_t hi) {
    blend64_job_t* j = (blend64_job_t*)arg;
    for (uint64_t t = lo; t < hi; ++t)
        blend_tile_f64(j->sp, j->order + j->tstart[t], (uint32_t)(j->tstart[t + 1] - j->tstart[t]),
                       (uint32_t)(t % j->gx), (uint32_t)(t / j->gx), j->W, j->H, j->image, NULL);
}

typedef struct {
    const float* gauss; const uint32_t* ent_idx; const double* VMs; const double* PVMs; const double* V;
    double fx, fy; uint32_t W, H, deg; splat_t_f64* out;
} p64_job_t;

/* Returns the FP64 pair count; image is W*H*3 doubles. */
uint64_t tpo_render_f64(const float* gauss, uint32_t n, const uint32_t* ent_idx, const float* models, uint32_t n_ent,
                        const float* view, const float* proj_view, const float* focal_ndc, uint32_t W, uint32_t H,
                        uint32_t sh_degree, double* image, int nthreads) {
    const uint32_t ne = n_ent ? n_ent : 1u;
    double V[16], PV[16];
    for (int i = 0; i < 16; ++i) { V[i] = view[i]; PV[i] = proj_view[i]; }
    double* VMs = (double*)malloc(sizeof(double) * 16 * ne);
    double* PVMs = (double*)malloc(sizeof(double) * 16 * ne);
    for (uint32_t e = 0; e < ne; ++e) {
        double M[16];
        for (int i = 0; i < 16; ++i) M[i] = models ? (double)models[16 * e + i] : ((i % 5) == 0 ? 1.0 : 0.0);
        mat4_mul_plain_f64(V, M, VMs + 16 * e);
        mat4_mul_plain_f64(PV, M, PVMs + 16 * e);
    }
    splat_t_f64* sp = (splat_t_f64*)malloc(sizeof(splat_t_f64) * (n ? n : 1));
    project_job_t_f64 job = {gauss, n, models ? ent_idx : NULL, VMs, PVMs, V,
                             (0.5 * (double)W) * (double)focal_ndc[0], (0.5 * (double)H) * (double)focal_ndc[1], W, H,
                             sh_degree > 3 ? 3 : sh_degree, sp};
    parallel_for(project_range_f64, &job, n, nthreads);
    const uint32_t gx = (W + 15u) / 16u, gy = (H + 15u) / 16u, ntiles = gx * gy;
    uint64_t P = 0;
    for (uint32_t i = 0; i < n; ++i) P += sp[i].tiles;
    pair64_t* pairs = (pair64_t*)malloc(sizeof(pair64_t) * (P ? P : 1));
    uint64_t k = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (sp[i].tiles == 0) continue;
        for (int32_t y = sp[i].rect[1]; y < sp[i].rect[3]; ++y)
            for (int32_t x = sp[i].rect[0]; x < sp[i].rect[2]; ++x) {
                pairs[k].tile = (uint32_t)y * gx + (uint32_t)x;
                pairs[k].depth = sp[i].texel[2];
                pairs[k].idx = i;
                ++k;
            }
    }
    qsort(pairs, P, sizeof(pair64_t), cmp_pair64);
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (P ? P : 1));
    uint64_t* tstart = (uint64_t*)calloc((size_t)ntiles + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < P; ++i) { order[i] = pairs[i].idx; tstart[pairs[i].tile + 1]++; }
    for (uint32_t t = 0; t < ntiles; ++t) tstart[t + 1] += tstart[t];
    blend64_job_t bj = {sp, order, tstart, W, H, gx, image};
    parallel_for(blend64_range, &bj, ntiles, nthreads);
    free(VMs); free(PVMs); free(sp); free(pairs); free(order); free(tstart);
    return P;
}
