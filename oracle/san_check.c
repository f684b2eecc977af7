/*
 * san_check.c -- drives every entry point of cpu_ref.c (FP32 and FP64
 * builds) over small 2D, 3D and degenerate shapes, for a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle sanitize).
 * TEST INFRASTRUCTURE ONLY: run by tests/test_oracle_sanitize.py.
 *
 * Any out-of-bounds access, use after free, leak, signed overflow or other
 * undefined behaviour aborts the run with the sanitizer's report; a clean
 * run prints the number of calls made and exits 0.
 */
#include <stdio.h>
#include <stdlib.h>

#define DECLARE(R, S)                                                                                            \
    R cr_rof_##S(const R *, R *, R *, const R *, int, int, R, R, int, int, int);                                 \
    R cr_fgp_##S(const R *, R *, R *, R, int, R, int, int, int, int, int);                                       \
    R cr_dtv_##S(const R *, const R *, R *, R *, R, int, R, R, int, int, int, int, int);                         \
    R cr_pd_##S(const R *, R *, R *, R, int, R, R, int, int, int, int, int);                                     \
    R cr_sb_##S(const R *, R *, R *, R, int, R, int, int, int, int);                                             \
    R cr_tgv_##S(const R *, R *, R *, R, R, R, int, R, R, int, int, int);                                        \
    R cr_llt_##S(const R *, R *, R *, R, R, int, R, R, int, int, int);                                           \
    R cr_ndf_##S(const R *, R *, R *, R, R, int, R, int, R, int, int, int);                                      \
    R cr_ndf_inplace_##S(const R *, R *, R *, R, R, int, R, int, R, int, int, int);                              \
    R cr_d4_##S(const R *, R *, R *, R, R, int, R, R, int, int, int);                                            \
    void cr_set_forced_stop_##S(int);

DECLARE(float, f32)
DECLARE(double, f64)
double cr_tv_energy_f64(const float *, const float *, double, int, int, int, int);

static int calls = 0;

/* Every solver on one shape: fixed counts and tolerance stops, iso and
 * aniso, nonneg on and off, every NDF penalty, a forced stop. */
#define RUN_ALL(R, S)                                                                                            \
    static void run_all_##S(int nx, int ny, int nz) {                                                            \
        const size_t n = (size_t)nx * ny * nz;                                                                   \
        R *in = malloc(n * sizeof(R)), *ref = malloc(n * sizeof(R)), *out = malloc(n * sizeof(R));              \
        R *lam = malloc(n * sizeof(R));                                                                          \
        R info[2];                                                                                               \
        unsigned s = 12345u + (unsigned)n;                                                                       \
        for (size_t i = 0; i < n; ++i) {                                                                         \
            s = s * 1664525u + 1013904223u;                                                                      \
            in[i] = (R)((s >> 8) & 0xffff) / (R)65536;                                                           \
            ref[i] = in[(i * 7) % n];                                                                            \
            lam[i] = (R)0.02 + (R)0.01 * in[i];                                                                  \
        }                                                                                                        \
        const R l = (R)0.03;                                                                                     \
        for (int tol = 0; tol < 2; ++tol) {                                                                      \
            const R eps = tol ? (R)1e-3 : (R)0;                                                                  \
            for (int m = 0; m < 2; ++m) {                                                                        \
                cr_fgp_##S(in, out, info, l, 12, eps, m, m ^ 1, nx, ny, nz); ++calls;                            \
                cr_pd_##S(in, out, info, l, 12, eps, (R)12, m, m, nx, ny, nz); ++calls;                          \
                cr_sb_##S(in, out, info, l, 12, eps, m, nx, ny, nz); ++calls;                                    \
                cr_dtv_##S(in, ref, out, info, l, 12, eps, (R)0.01, m, m ^ 1, nx, ny, nz); ++calls;              \
            }                                                                                                    \
            cr_rof_##S(in, out, info, &l, 0, 12, (R)0.002, eps, nx, ny, nz); ++calls;                            \
            cr_rof_##S(in, out, info, lam, 1, 12, (R)0.002, eps, nx, ny, nz); ++calls;                           \
            cr_tgv_##S(in, out, info, l, (R)1, (R)2, 12, (R)12, eps, nx, ny, nz); ++calls;                       \
            cr_llt_##S(in, out, info, l, (R)0.01, 12, (R)0.002, eps, nx, ny, nz); ++calls;                       \
            for (int pen = 1; pen <= 5; ++pen) {                                                                 \
                cr_ndf_##S(in, out, info, l, (R)0.05, 12, (R)0.01, pen, eps, nx, ny, nz); ++calls;               \
                cr_ndf_inplace_##S(in, out, info, l, (R)0.05, 12, (R)0.01, pen, eps, nx, ny, nz); ++calls;       \
            }                                                                                                    \
            cr_ndf_##S(in, out, info, l, (R)0, 12, (R)0.01, 1, eps, nx, ny, nz); ++calls;                        \
            cr_d4_##S(in, out, info, l, (R)0.05, 12, (R)1e-3, eps, nx, ny, nz); ++calls;                         \
        }                                                                                                        \
        cr_set_forced_stop_##S(5);                                                                               \
        cr_fgp_##S(in, out, info, l, 12, (R)1e-30, 0, 1, nx, ny, nz); ++calls;                                   \
        cr_tgv_##S(in, out, info, l, (R)1, (R)2, 12, (R)12, (R)1e-30, nx, ny, nz); ++calls;                      \
        cr_set_forced_stop_##S(-1);                                                                              \
        free(in); free(ref); free(out); free(lam);                                                               \
    }

RUN_ALL(float, f32)
RUN_ALL(double, f64)

int main(void) {
    static const int shapes[][3] = {{13, 17, 1}, {66, 3, 1}, {9, 7, 5}, {66, 1, 9}, {1, 9, 6}, {2, 2, 2}, {70, 5, 3}};
    for (size_t k = 0; k < sizeof shapes / sizeof shapes[0]; ++k) {
        run_all_f32(shapes[k][0], shapes[k][1], shapes[k][2]);
        run_all_f64(shapes[k][0], shapes[k][1], shapes[k][2]);
        const size_t n = (size_t)shapes[k][0] * shapes[k][1] * shapes[k][2];
        float *u = calloc(n, sizeof(float)), *u0 = calloc(n, sizeof(float));
        for (size_t i = 0; i < n; ++i) u[i] = (float)(i % 7);
        (void)cr_tv_energy_f64(u, u0, 0.1, 1, shapes[k][0], shapes[k][1], shapes[k][2]);
        (void)cr_tv_energy_f64(u, u0, 0.1, 2, shapes[k][0], shapes[k][1], shapes[k][2]);
        calls += 2;
        free(u);
        free(u0);
    }
    printf("sanitizer check: %d oracle calls clean\n", calls);
    return 0;
}
