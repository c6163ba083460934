/*
 * psm_oracle.c -- CPU restatement of the reference diffraction-summation path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * poststackmig driver, the Python host mirror) links, loads or calls this
 * file.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, and only as the checker.
 *
 * What it restates (reference = sergeevsn/poststackmig3d):
 *   psmo_kernel_inline_f32   MigrationKernel::kirchhoffKernel3DWithPadding,
 *                            src/migration_kernel.cpp:17-383, OpenMP build
 *                            (nested team inactive => one thread walks the
 *                            crosslines of a call in ascending order, so the
 *                            per-sample caches of :64-72 are "sticky",
 *                            :115-152).
 *   psmo_kernel_inline_f64   the same loop nest evaluated in double precision
 *                            (accuracy yardstick; not a reference function).
 *   psmo_migrate_volume_*    the driver loop of src/main.cpp:140-217,254-383
 *                            (padded output labels, nearest-real-inline
 *                            mapping, aperture window of
 *                            migration_kernel.cpp:385-468, velocity slice of
 *                            :470-511) over an in-memory cube.
 *   psmo_estimate_max_velocity  main.cpp:22-64 on a velocity cube.
 *   psmo_adjoint_volume_f64  transpose of psmo_migrate_volume_f64 (no
 *                            reference counterpart; extension).
 *   psmo_count_pairs         the pair counts P_ap / P_g of the work unit.
 *
 * Pinning (see tests/test_oracle.py): golden vectors written by the reference
 * build in tests/golden/, the reference library oracle/_ref (when built),
 * closed forms (single spike, constant trace count, r=0 term, t0=0 row),
 * linearity, mirror symmetry, and brute-force double evaluation.
 *
 * Arithmetic: compiled with -ffp-contract=off so that every fused
 * multiply-add below is written out explicitly with fmaf(); the reference's
 * -ffast-math build contracts r^2 = dy*dy + dx^2 and tau^2 = a*r^2 + t0^2 into
 * FMAs (checked against oracle/_ref by tests/test_oracle.py).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PSMO_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Per-call kernel, single precision (migration_kernel.cpp:17-383).          */
/* ------------------------------------------------------------------------ */
PSMO_API int psmo_kernel_inline_f32(
    const float* in, size_t n_buf_il, size_t n_buf_xl, size_t n_t,
    const float* vel, float* out, size_t n_xl,
    const int32_t* il_pos, size_t n_il_pos, int32_t cur_il_pos,
    int32_t cur_xl_start, float dt, const float* t0_times,
    float tan_theta_sq, float dx, float dy, int amp)
{
    if (n_xl == 0 || n_t == 0) return 0;                    /* :35-37 */
    const size_t stride_il = n_buf_xl * n_t;                 /* :40-41 */
    const float inv_dt = 1.0f / dt;                          /* :44 */
    const float inv_dy = 1.0f / dy;                          /* :45 */
    const int32_t cpos = (cur_il_pos >= 0 && (size_t)cur_il_pos < n_il_pos)
                             ? il_pos[cur_il_pos] : 0;       /* :48-49 */
    float* dx2 = (float*)malloc(sizeof(float) * (n_buf_il ? n_buf_il : 1));
    for (size_t b = 0; b < n_buf_il; ++b) {                  /* :51-56 */
        float d = (float)(il_pos[b] - cpos) * dx;
        dx2[b] = d * d;
    }
    /* caches live for the whole call (:64-72) */
    float* c_t0sq = (float*)calloc(n_t, sizeof(float));
    float* c_v = (float*)calloc(n_t, sizeof(float));
    float* c_a = (float*)calloc(n_t, sizeof(float));
    float* c_R2 = (float*)calloc(n_t, sizeof(float));
    int* c_lo = (int*)calloc(n_t, sizeof(int));
    int* c_hi = (int*)calloc(n_t, sizeof(int));
    unsigned char* c_valid = (unsigned char*)calloc(n_t, 1);

    for (size_t j = 0; j < n_xl; ++j) {                      /* :75 */
        const int32_t kj = cur_xl_start + (int32_t)j;         /* :77 */
        float* o = out + j * n_t;
        if (kj < 0 || kj >= (int32_t)n_buf_xl) {             /* :79-85 */
            for (size_t s = 0; s < n_t; ++s) o[s] = 0.0f;
            continue;
        }
        for (size_t s = 0; s < n_t; ++s) {                   /* :115-144 */
            const float t0 = t0_times[s];
            const float v = vel[(size_t)kj * n_t + s];
            if (v >= 1e-5f && t0 >= 1e-5f) {
                c_valid[s] = 1;
                c_t0sq[s] = t0 * t0;
                c_v[s] = v;
                const float inv_v2 = 1.0f / (v * v + 1e-10f);
                c_a[s] = 4.0f * inv_v2;
                const float vth = v * t0 * 0.5f;
                c_R2[s] = vth * vth * tan_theta_sq;
                const float max_r = sqrtf(c_R2[s]);
                const int w = (int)(max_r * inv_dy) + 1;
                c_lo[s] = kj - w > 0 ? kj - w : 0;
                c_hi[s] = kj + w + 1 < (int)n_buf_xl ? kj + w + 1 : (int)n_buf_xl;
            }
        }
        for (size_t s = 0; s < n_t; ++s) {                   /* :147-376 */
            if (!c_valid[s]) { o[s] = 0.0f; continue; }
            const float t0sq = c_t0sq[s], v = c_v[s], a = c_a[s], R2 = c_R2[s];
            const float v2 = v * v;
            float acc = 0.0f;
            for (size_t b = 0; b < n_buf_il; ++b) {          /* :163-168 */
                const float d2 = dx2[b];
                if (d2 > R2) continue;
                const float* row = in + b * stride_il;
                for (int k = c_lo[s]; k < c_hi[s]; ++k) {    /* :188-320 */
                    const float dyv = (float)(k - kj) * dy;
                    const float r2 = fmaf(dyv, dyv, d2);
                    if (r2 > R2) continue;
                    const float tau = sqrtf(fmaf(a, r2, t0sq));
                    const float q = tau * inv_dt;
                    const int i = (int)q;
                    if (i < (int)n_t - 1) {
                        const float f = q - (float)i;
                        const float* tr = row + (size_t)k * n_t;
                        float val = fmaf(f, tr[i + 1] - tr[i], tr[i]);  /* :544-550 */
                        if (amp) val *= 1.0f / fmaf(tau, v2, 1e-10f);    /* :265-272 */
                        acc += val;
                    }
                }
            }
            o[s] = acc;                                      /* :375 */
        }
    }
    free(dx2); free(c_t0sq); free(c_v); free(c_a); free(c_R2);
    free(c_lo); free(c_hi); free(c_valid);
    return 0;
}

/* The float32 decision of migration_kernel.cpp:298 (i < n_t - 1) for one
 * pair, so that the double-precision oracles sum the same pair set; returns
 * the double-precision (i, f), clamped to the last interval when the two
 * precisions straddle the bound. */
static int bound_f32(float vf, float t0f, float r2f, float inv_dtf, int n_t,
                     double q, int* i_out, double* f_out)
{
    const float af = 4.0f * (1.0f / (vf * vf + 1e-10f));
    const float qf = sqrtf(fmaf(af, r2f, t0f * t0f)) * inv_dtf;
    if ((int)qf >= n_t - 1) return 0;
    int i = (int)q;
    double f = q - (double)i;
    if (i > n_t - 2) { i = n_t - 2; f = 1.0; }
    *i_out = i;
    *f_out = f;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Per-call kernel, double precision (accuracy yardstick, not a reference   */
/* function).  The discrete decisions -- sample validity, the aperture mask */
/* r^2 <= R^2, the crossline window and the dx^2 > R^2 row skip -- are made */
/* in float exactly as psmo_kernel_inline_f32 makes them, so that both      */
/* oracles sum the same pair set; the travel time, interpolation, weight    */
/* and accumulation are evaluated in double.  (Evaluating the mask in       */
/* double moves exact ties such as r = R on integer grids at 45 degrees,    */
/* which is a change of operator, not an arithmetic error.)                 */
/* ------------------------------------------------------------------------ */
PSMO_API int psmo_kernel_inline_f64(
    const float* in, size_t n_buf_il, size_t n_buf_xl, size_t n_t,
    const float* vel, double* out, size_t n_xl,
    const int32_t* il_pos, size_t n_il_pos, int32_t cur_il_pos,
    int32_t cur_xl_start, double dt, const float* t0_times, float tan_theta_sq,
    double dx, double dy, int amp)
{
    if (n_xl == 0 || n_t == 0) return 0;
    const size_t stride_il = n_buf_xl * n_t;
    const float dxf = (float)dx, dyf = (float)dy, inv_dy = 1.0f / dyf;
    const float inv_dtf = 1.0f / (float)dt;
    const int32_t cpos = (cur_il_pos >= 0 && (size_t)cur_il_pos < n_il_pos)
                             ? il_pos[cur_il_pos] : 0;
    float* dx2f = (float*)malloc(sizeof(float) * (n_buf_il ? n_buf_il : 1));
    double* dx2 = (double*)malloc(sizeof(double) * (n_buf_il ? n_buf_il : 1));
    for (size_t b = 0; b < n_buf_il; ++b) {
        float d = (float)(il_pos[b] - cpos) * dxf;
        dx2f[b] = d * d;
        double dd = (double)(il_pos[b] - cpos) * dx;
        dx2[b] = dd * dd;
    }
    double* c_v = (double*)calloc(n_t, sizeof(double));
    float* c_R2 = (float*)calloc(n_t, sizeof(float));
    int* c_lo = (int*)calloc(n_t, sizeof(int));
    int* c_hi = (int*)calloc(n_t, sizeof(int));
    unsigned char* c_valid = (unsigned char*)calloc(n_t, 1);
    for (size_t j = 0; j < n_xl; ++j) {
        const int32_t kj = cur_xl_start + (int32_t)j;
        double* o = out + j * n_t;
        if (kj < 0 || kj >= (int32_t)n_buf_xl) {
            for (size_t s = 0; s < n_t; ++s) o[s] = 0.0;
            continue;
        }
        for (size_t s = 0; s < n_t; ++s) {
            const float t0 = t0_times[s];
            const float v = vel[(size_t)kj * n_t + s];
            if (v >= 1e-5f && t0 >= 1e-5f) {
                c_valid[s] = 1;
                c_v[s] = (double)v;
                const float vth = v * t0 * 0.5f;
                c_R2[s] = vth * vth * tan_theta_sq;
                const int w = (int)(sqrtf(c_R2[s]) * inv_dy) + 1;
                c_lo[s] = kj - w > 0 ? kj - w : 0;
                c_hi[s] = kj + w + 1 < (int)n_buf_xl ? kj + w + 1 : (int)n_buf_xl;
            }
        }
        for (size_t s = 0; s < n_t; ++s) {
            if (!c_valid[s]) { o[s] = 0.0; continue; }
            const double t0 = (double)s * dt, v = c_v[s], a = 4.0 / (v * v);
            const float R2 = c_R2[s];
            double acc = 0.0;
            for (size_t b = 0; b < n_buf_il; ++b) {
                if (dx2f[b] > R2) continue;
                const float* row = in + b * stride_il;
                for (int k = c_lo[s]; k < c_hi[s]; ++k) {
                    const float dyv = (float)(k - kj) * dyf;
                    const float r2f = fmaf(dyv, dyv, dx2f[b]);
                    if (r2f > R2) continue;
                    const double dyd = (double)(k - kj) * dy;
                    const double r2 = dyd * dyd + dx2[b];
                    const double tau = sqrt(t0 * t0 + a * r2);
                    int i;
                    double f;
                    if (bound_f32((float)v, t0_times[s], r2f, inv_dtf, (int)n_t, tau / dt, &i, &f)) {
                        const float* tr = row + (size_t)k * n_t;
                        double val = (double)tr[i] + f * ((double)tr[i + 1] - (double)tr[i]);
                        if (amp) val /= tau * v * v;
                        acc += val;
                    }
                }
            }
            o[s] = acc;
        }
    }
    free(dx2f); free(dx2); free(c_v); free(c_R2);
    free(c_lo); free(c_hi); free(c_valid);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Driver geometry (main.cpp:22-64, 140-217).                                */
/* ------------------------------------------------------------------------ */

/* main.cpp:22-64 on a velocity cube [n_il][n_xl][n_t] laid out on the input
 * grid: probes il,xl in {0, n/2, n-1} at the last sample; 3000 if < 100. */
PSMO_API float psmo_estimate_max_velocity(const float* vel, size_t n_il,
                                          size_t n_xl, size_t n_t)
{
    size_t ils[3] = {0, n_il / 2, n_il - 1}, xls[3] = {0, n_xl / 2, n_xl - 1};
    size_t nil = n_il > 1 ? 3 : 1, nxl = n_xl > 1 ? 3 : 1;
    float vmax = 0.0f;
    for (size_t a = 0; a < nil; ++a)
        for (size_t b = 0; b < nxl; ++b) {
            float v = vel[(ils[a] * n_xl + xls[b]) * n_t + (n_t - 1)];
            if (v > vmax) vmax = v;
        }
    if (vmax < 100.0f) vmax = 3000.0f;
    return vmax;
}

/* main.cpp:188-202: tan^2 of the aperture and the inline window half width. */
PSMO_API void psmo_aperture(double angle_deg, float vmax, size_t n_t, double dt,
                            double il_step, float* tan_theta_sq, int* max_ap_il)
{
    const float tan_theta = (float)tan(angle_deg * 3.14159265358979323846 / 180.0);
    const float t2 = tan_theta * tan_theta;
    const double max_t = (double)(n_t - 1) * dt;
    const float radius = (vmax * (float)max_t / 2.0f) * sqrtf(t2);
    *tan_theta_sq = t2;
    *max_ap_il = (int)((double)radius / il_step) + 1;
}

/* main.cpp:144-172: output labels extended by pad steps of the first-two-label
 * step.  Returns the number of labels written (out may be NULL to count). */
PSMO_API size_t psmo_output_labels(const int32_t* labels, size_t n, int pad,
                                   int32_t* out)
{
    if (pad > 0 && n > 1) {
        const int32_t step = labels[1] - labels[0];
        const int32_t first = labels[0] - pad * step;
        const int32_t last = labels[n - 1] + pad * step;
        size_t c = 0;
        if (step <= 0) return 0;
        for (int32_t l = first; l <= last; l += step) {
            if (out) out[c] = l;
            ++c;
        }
        return c;
    }
    if (out) memcpy(out, labels, n * sizeof(int32_t));
    return n;
}

/* main.cpp:258-288: output inline label -> index of the real inline used. */
PSMO_API size_t psmo_map_output_inline(const int32_t* labels, size_t n,
                                       int32_t out_label)
{
    for (size_t i = 0; i < n; ++i)
        if (labels[i] == out_label) return i;
    if (out_label < labels[0]) return 0;
    if (out_label > labels[n - 1]) return n - 1;
    size_t best = 0;
    int32_t md = abs(out_label - labels[0]);
    for (size_t i = 1; i < n; ++i) {
        int32_t d = abs(out_label - labels[i]);
        if (d < md) { md = d; best = i; }
    }
    return best;
}

typedef struct {
    size_t n_il, n_xl, n_t;
    double dt, il_step, xl_step, angle_deg;
    int pad_il, pad_xl, amp;
} psmo_volume;

/* Builds the per-output-inline buffer exactly as
 * MigrationKernel::prepareInputBufferWithPadding (migration_kernel.cpp:385-468)
 * and prepareVelocitySliceWithPadding (:470-511), then finds the current
 * buffer position (main.cpp:329-352). */
static void build_window(const float* data, const float* vel, const psmo_volume* g,
                         size_t c, int max_ap_il, float* buf, float* vslice,
                         int32_t* il_pos, size_t* n_buf_il, int32_t* cur)
{
    const size_t n_il = g->n_il, n_xl = g->n_xl, n_t = g->n_t;
    const int padx = g->pad_xl;
    int s_req = (int)c - max_ap_il; if (s_req < 0) s_req = 0;
    int e_req = (int)c + max_ap_il + 1; if (e_req > (int)n_il) e_req = (int)n_il;
    int s_buf = s_req - g->pad_il; if (s_buf < 0) s_buf = 0;
    int e_buf = e_req + g->pad_il; if (e_buf > (int)n_il) e_buf = (int)n_il;
    const size_t K = n_xl + 2 * (size_t)padx;
    *n_buf_il = (size_t)(e_buf - s_buf);
    memset(buf, 0, sizeof(float) * (*n_buf_il) * K * n_t);
    for (int b = s_buf; b < e_buf; ++b) {
        il_pos[b - s_buf] = b;
        for (size_t x = 0; x < n_xl; ++x)
            memcpy(buf + ((size_t)(b - s_buf) * K + padx + x) * n_t,
                   data + ((size_t)b * n_xl + x) * n_t, n_t * sizeof(float));
    }
    memset(vslice, 0, sizeof(float) * K * n_t);
    for (size_t x = 0; x < n_xl; ++x)
        memcpy(vslice + (padx + x) * n_t, vel + (c * n_xl + x) * n_t, n_t * sizeof(float));
    *cur = -1;
    for (size_t i = 0; i < *n_buf_il; ++i)
        if (il_pos[i] == (int32_t)c) { *cur = (int32_t)i; break; }
    if (*cur < 0) {
        int32_t best = 0, md = abs(il_pos[0] - (int32_t)c);
        for (size_t i = 1; i < *n_buf_il; ++i) {
            int32_t d = abs(il_pos[i] - (int32_t)c);
            if (d < md) { md = d; best = (int32_t)i; }
        }
        *cur = best;
    }
}

/*
 * Whole-volume migration (main.cpp:254-383) on an in-memory cube.
 *   data, vel : [n_il][n_xl][n_t] on the input grid (velocity of the image
 *               point's inline/crossline; constant velocity = filled cube)
 *   il_labels, xl_labels : sorted labels of the input grid (NULL -> 1..n)
 *   out       : [n_out_il][n_out_xl][n_t]; only the rows listed in `rows`
 *               (NULL -> all) are written.
 * precision 0 = f32 restatement (out is float*), 1 = f64 (out is double*).
 */
static int migrate_volume(const float* data, const float* vel, const psmo_volume* g,
                          const int32_t* il_labels, const int32_t* xl_labels,
                          void* out, size_t n_out_il, size_t n_out_xl,
                          const int32_t* rows, size_t n_rows, int nthreads,
                          int precision)
{
    const size_t n_il = g->n_il, n_xl = g->n_xl, n_t = g->n_t;
    int32_t* ill = (int32_t*)malloc(sizeof(int32_t) * n_il);
    for (size_t i = 0; i < n_il; ++i) ill[i] = il_labels ? il_labels[i] : (int32_t)(i + 1);
    (void)xl_labels;
    size_t nol = psmo_output_labels(ill, n_il, g->pad_il, NULL);
    int32_t* oll = (int32_t*)malloc(sizeof(int32_t) * (nol ? nol : 1));
    psmo_output_labels(ill, n_il, g->pad_il, oll);
    if (nol != n_out_il) { free(ill); free(oll); return -1; }

    const float vmax = psmo_estimate_max_velocity(vel, n_il, n_xl, n_t);
    float t2; int ap;
    psmo_aperture(g->angle_deg, vmax, n_t, g->dt, g->il_step, &t2, &ap);
    float* t0 = (float*)malloc(sizeof(float) * n_t);
    for (size_t i = 0; i < n_t; ++i) t0[i] = (float)((double)i * g->dt);  /* main.cpp:214-217 */
    const size_t K = n_xl + 2 * (size_t)g->pad_xl;
    const size_t nr = rows ? n_rows : n_out_il;
    int rc = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t r = 0; r < nr; ++r) {
        const size_t io = rows ? (size_t)rows[r] : r;
        const size_t c = psmo_map_output_inline(ill, n_il, oll[io]);
        size_t wmax = 2 * (size_t)(ap + g->pad_il) + 1;
        if (wmax > n_il) wmax = n_il;
        float* buf = (float*)malloc(sizeof(float) * wmax * K * n_t);
        float* vs = (float*)malloc(sizeof(float) * K * n_t);
        int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * wmax);
        size_t nb; int32_t cur;
        build_window(data, vel, g, c, ap, buf, vs, pos, &nb, &cur);
        if (precision == 0) {
            psmo_kernel_inline_f32(buf, nb, K, n_t, vs, (float*)out + io * n_out_xl * n_t,
                                   n_out_xl, pos, nb, cur, 0, (float)g->dt, t0, t2,
                                   (float)g->il_step, (float)g->xl_step, g->amp);
        } else {
            psmo_kernel_inline_f64(buf, nb, K, n_t, vs, (double*)out + io * n_out_xl * n_t,
                                   n_out_xl, pos, nb, cur, 0, g->dt, t0, t2,
                                   g->il_step, g->xl_step, g->amp);
        }
        free(buf); free(vs); free(pos);
    }
    free(ill); free(oll); free(t0);
    return rc;
}

PSMO_API int psmo_migrate_volume_f32(const float* data, const float* vel,
                                     const psmo_volume* g, const int32_t* il_labels,
                                     const int32_t* xl_labels, float* out,
                                     size_t n_out_il, size_t n_out_xl,
                                     const int32_t* rows, size_t n_rows, int nthreads)
{
    return migrate_volume(data, vel, g, il_labels, xl_labels, out, n_out_il, n_out_xl,
                          rows, n_rows, nthreads, 0);
}

PSMO_API int psmo_migrate_volume_f64(const float* data, const float* vel,
                                     const psmo_volume* g, const int32_t* il_labels,
                                     const int32_t* xl_labels, double* out,
                                     size_t n_out_il, size_t n_out_xl,
                                     const int32_t* rows, size_t n_rows, int nthreads)
{
    return migrate_volume(data, vel, g, il_labels, xl_labels, out, n_out_il, n_out_xl,
                          rows, n_rows, nthreads, 1);
}

/* ------------------------------------------------------------------------ */
/* Adjoint (demigration) of psmo_migrate_volume_f64 -- extension, no        */
/* reference counterpart.  Regular labels only (1..n).  Walks the same pair */
/* set as the forward operator (float mask decisions, sticky parameter      */
/* source) and scatters d[b][k][i] += w (1-f) m, d[b][k][i+1] += w f m.     */
/* Crossline padding traces are zero inputs of the forward operator, so     */
/* their adjoint contributions are dropped.                                 */
/* ------------------------------------------------------------------------ */
PSMO_API int psmo_adjoint_volume_f64(const double* image, const float* vel,
                                     const psmo_volume* g, double* data_out,
                                     size_t n_out_il, size_t n_out_xl)
{
    const size_t n_il = g->n_il, n_xl = g->n_xl, n_t = g->n_t;
    const int padx = g->pad_xl;
    const size_t K = n_xl + 2 * (size_t)padx;
    int32_t* ill = (int32_t*)malloc(sizeof(int32_t) * n_il);
    for (size_t i = 0; i < n_il; ++i) ill[i] = (int32_t)(i + 1);
    size_t nol = psmo_output_labels(ill, n_il, g->pad_il, NULL);
    int32_t* oll = (int32_t*)malloc(sizeof(int32_t) * (nol ? nol : 1));
    psmo_output_labels(ill, n_il, g->pad_il, oll);
    if (nol != n_out_il || n_out_xl != K) { free(ill); free(oll); return -1; }
    const float vmax = psmo_estimate_max_velocity(vel, n_il, n_xl, n_t);
    float T2; int ap;
    psmo_aperture(g->angle_deg, vmax, n_t, g->dt, g->il_step, &T2, &ap);
    const double dt = g->dt, dx = g->il_step, dy = g->xl_step;
    const float dxf = (float)dx, dyf = (float)dy, inv_dy = 1.0f / dyf;
    const float inv_dtf = 1.0f / (float)dt;
    memset(data_out, 0, sizeof(double) * n_il * n_xl * n_t);
    for (size_t io = 0; io < n_out_il; ++io) {
        const size_t c = psmo_map_output_inline(ill, n_il, oll[io]);
        int s_req = (int)c - ap; if (s_req < 0) s_req = 0;
        int e_req = (int)c + ap + 1; if (e_req > (int)n_il) e_req = (int)n_il;
        int b0 = s_req - g->pad_il; if (b0 < 0) b0 = 0;
        int b1 = e_req + g->pad_il; if (b1 > (int)n_il) b1 = (int)n_il;
        for (size_t s = 0; s < n_t; ++s) {
            const float t0f = (float)((double)s * dt);
            const double t0 = (double)s * dt;
            int src = -1;
            for (size_t j = 0; j < n_out_xl; ++j) {
                const int kj = (int)j;
                const int vx = kj - padx;
                if (vx >= 0 && vx < (int)n_xl) {
                    const float v = vel[(c * n_xl + vx) * n_t + s];
                    if (v >= 1e-5f && t0f >= 1e-5f) src = kj;
                }
                if (src < 0) continue;
                const double m = image[(io * n_out_xl + j) * n_t + s];
                if (m == 0.0) continue;
                const float vf = vel[(c * n_xl + (src - padx)) * n_t + s];
                const float vth = vf * t0f * 0.5f;
                const float R2 = vth * vth * T2;
                const int w = (int)(sqrtf(R2) * inv_dy) + 1;
                const int lo = src - w > 0 ? src - w : 0;
                const int hi = src + w + 1 < (int)K ? src + w + 1 : (int)K;
                const double v = (double)vf, a = 4.0 / (v * v);
                for (int b = b0; b < b1; ++b) {
                    const float dxv = (float)(b - (int)c) * dxf;
                    const float d2f = dxv * dxv;
                    if (d2f > R2) continue;
                    const double dxd = (double)(b - (int)c) * dx;
                    for (int k = lo; k < hi; ++k) {
                        const int x = k - padx;
                        if (x < 0 || x >= (int)n_xl) continue;
                        const float dyv = (float)(k - kj) * dyf;
                        if (fmaf(dyv, dyv, d2f) > R2) continue;
                        const double dyd = (double)(k - kj) * dy;
                        const double r2 = dyd * dyd + dxd * dxd;
                        const double tau = sqrt(t0 * t0 + a * r2);
                        int i;
                        double f;
                        if (!bound_f32(vf, t0f, fmaf(dyv, dyv, d2f), inv_dtf, (int)n_t, tau / dt,
                                       &i, &f))
                            continue;
                        double wgt = m;
                        if (g->amp) wgt /= tau * v * v;
                        double* tr = data_out + ((size_t)b * n_xl + x) * n_t;
                        tr[i] += wgt * (1.0 - f);
                        tr[i + 1] += wgt * f;
                    }
                }
            }
        }
    }
    free(ill); free(oll);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Pair counts (work unit), brute force over the reference loop nest.        */
/*   p_ap: (output sample, aperture trace) pairs that pass the aperture mask */
/*         of migration_kernel.cpp:166,286 (r^2 <= R^2), over every output   */
/*         sample the reference evaluates (sticky samples included), with    */
/*         the aperture trace restricted to real input traces (crossline     */
/*         padding traces are zero and contribute nothing)                   */
/*   p_g : the subset that also has i < n_t - 1 (:298), i.e. gathers        */
/* Regular labels (1..n); f32 arithmetic as psmo_kernel_inline_f32.          */
/* ------------------------------------------------------------------------ */
PSMO_API int psmo_count_pairs(const float* vel, const psmo_volume* g,
                              const int32_t* rows, size_t n_rows,
                              double* p_ap, double* p_g, int nthreads)
{
    const size_t n_il = g->n_il, n_xl = g->n_xl, n_t = g->n_t;
    const int padx = g->pad_xl;
    const size_t K = n_xl + 2 * (size_t)padx;
    int32_t* ill = (int32_t*)malloc(sizeof(int32_t) * n_il);
    for (size_t i = 0; i < n_il; ++i) ill[i] = (int32_t)(i + 1);
    size_t nol = psmo_output_labels(ill, n_il, g->pad_il, NULL);
    int32_t* oll = (int32_t*)malloc(sizeof(int32_t) * (nol ? nol : 1));
    psmo_output_labels(ill, n_il, g->pad_il, oll);
    const float vmax = psmo_estimate_max_velocity(vel, n_il, n_xl, n_t);
    float T2; int ap;
    psmo_aperture(g->angle_deg, vmax, n_t, g->dt, g->il_step, &T2, &ap);
    const float dtf = (float)g->dt, inv_dt = 1.0f / dtf;
    const float dx = (float)g->il_step, dy = (float)g->xl_step, inv_dy = 1.0f / dy;
    const size_t nr = rows ? n_rows : nol;
    double tot_ap = 0.0, tot_g = 0.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : tot_ap, tot_g)
    for (size_t r = 0; r < nr; ++r) {
        const size_t io = rows ? (size_t)rows[r] : r;
        const size_t c = psmo_map_output_inline(ill, n_il, oll[io]);
        int s_req = (int)c - ap; if (s_req < 0) s_req = 0;
        int e_req = (int)c + ap + 1; if (e_req > (int)n_il) e_req = (int)n_il;
        int b0 = s_req - g->pad_il; if (b0 < 0) b0 = 0;
        int b1 = e_req + g->pad_il; if (b1 > (int)n_il) b1 = (int)n_il;
        for (size_t s = 0; s < n_t; ++s) {
            const float t0 = (float)((double)s * g->dt);
            int src = -1;
            for (size_t j = 0; j < K; ++j) {
                const int kj = (int)j, vx = kj - padx;
                if (vx >= 0 && vx < (int)n_xl) {
                    const float v = vel[(c * n_xl + vx) * n_t + s];
                    if (v >= 1e-5f && t0 >= 1e-5f) src = kj;
                }
                if (src < 0) continue;
                const float v = vel[(c * n_xl + (src - padx)) * n_t + s];
                const float a = 4.0f * (1.0f / (v * v + 1e-10f));
                const float vth = v * t0 * 0.5f;
                const float R2 = vth * vth * T2;
                const int w = (int)(sqrtf(R2) * inv_dy) + 1;
                const int lo = src - w > 0 ? src - w : 0;
                const int hi = src + w + 1 < (int)K ? src + w + 1 : (int)K;
                long long cap = 0, cg = 0;
                for (int b = b0; b < b1; ++b) {
                    const float dxv = (float)(b - (int)c) * dx;
                    const float d2 = dxv * dxv;
                    if (d2 > R2) continue;
                    for (int k = lo; k < hi; ++k) {
                        if (k < padx || k >= padx + (int)n_xl) continue;
                        const float dyv = (float)(k - kj) * dy;
                        const float r2 = fmaf(dyv, dyv, d2);
                        if (r2 > R2) continue;
                        ++cap;
                        const float tau = sqrtf(fmaf(a, r2, t0 * t0));
                        const int i = (int)(tau * inv_dt);
                        if (i < (int)n_t - 1) ++cg;
                    }
                }
                tot_ap += (double)cap;
                tot_g += (double)cg;
            }
        }
    }
    *p_ap = tot_ap;
    *p_g = tot_g;
    free(ill); free(oll);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Bounded sample of the volume loop (CPU baseline, port kind): for each    */
/* listed output inline, the f32 restatement computes only the output       */
/* crosslines [xl0, xl0 + m) (current_xl_start = xl0).  Regular labels.     */
/* out: [n_rows][m][n_t].                                                   */
/* ------------------------------------------------------------------------ */
PSMO_API int psmo_migrate_sample(const float* data, const float* vel, const psmo_volume* g,
                                 const int32_t* rows, size_t n_rows, int xl0, size_t m,
                                 float* out, int nthreads)
{
    const size_t n_il = g->n_il, n_xl = g->n_xl, n_t = g->n_t;
    int32_t* ill = (int32_t*)malloc(sizeof(int32_t) * n_il);
    for (size_t i = 0; i < n_il; ++i) ill[i] = (int32_t)(i + 1);
    size_t nol = psmo_output_labels(ill, n_il, g->pad_il, NULL);
    int32_t* oll = (int32_t*)malloc(sizeof(int32_t) * (nol ? nol : 1));
    psmo_output_labels(ill, n_il, g->pad_il, oll);
    const float vmax = psmo_estimate_max_velocity(vel, n_il, n_xl, n_t);
    float t2; int ap;
    psmo_aperture(g->angle_deg, vmax, n_t, g->dt, g->il_step, &t2, &ap);
    float* t0 = (float*)malloc(sizeof(float) * n_t);
    for (size_t i = 0; i < n_t; ++i) t0[i] = (float)((double)i * g->dt);
    const size_t K = n_xl + 2 * (size_t)g->pad_xl;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t r = 0; r < n_rows; ++r) {
        const size_t c = psmo_map_output_inline(ill, n_il, oll[rows[r]]);
        size_t wmax = 2 * (size_t)(ap + g->pad_il) + 1;
        if (wmax > n_il) wmax = n_il;
        float* buf = (float*)malloc(sizeof(float) * wmax * K * n_t);
        float* vs = (float*)malloc(sizeof(float) * K * n_t);
        int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * wmax);
        size_t nb; int32_t cur;
        build_window(data, vel, g, c, ap, buf, vs, pos, &nb, &cur);
        psmo_kernel_inline_f32(buf, nb, K, n_t, vs, out + r * m * n_t, m, pos, nb, cur, xl0,
                               (float)g->dt, t0, t2, (float)g->il_step, (float)g->xl_step,
                               g->amp);
        free(buf); free(vs); free(pos);
    }
    free(ill); free(oll); free(t0);
    return 0;
}
