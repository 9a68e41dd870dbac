/*
 * FP64 CPU restatement of idi's time-dependent simulation (decaying emitters).
 *
 * TEST INFRASTRUCTURE ONLY: tests/ call it to CHECK the HIP library; the product
 * path never links, loads or calls it.
 *
 * Follows (the code is written from scratch):
 *   idi/simulation/time.py:52-91     far field: q = r_p/|r_p| - z_hat, s = q . r_j (length),
 *                                    a = exp(-i (s k - phi)) (/|r_p| with "scale"),
 *                                    t = -s/c + T_j; pixel grid p (i - N/2), rotation M
 *   idi/simulation/simtime.cu:43-77  near field: s = |r_j - r_p| - |r_p| + z_j, t = s/c + T_j,
 *                                    a = exp(i (s k + phi)) (/|r_j - r_p| with "scale")
 *   idi/simulation/time.py:7-16      decaying sum over arrivals sorted by time:
 *                                    x_i = a_i + exp(-(t_i - t_{i-1})/tau) x_{i-1}
 *   idi/simulation/time.py:36-44     I = sum_{i<n} -tau/2 |x_i|^2 expm1(-2 (t_{i+1} - t_i)/tau)
 *                                        + tau/2 |x_n|^2
 * The sort is stable in emitter order for equal times (qsort on (t, j)).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <omp.h>

typedef struct {
    int64_t n0, n1;
    double pixelsize, detz, k, off0, off1;
    double M[9];
    int32_t rot, nf, scale, secondorder;
} oracle_geom;

typedef struct {
    double t;
    int64_t j;
} arrival;

static int by_time(const void *a, const void *b)
{
    const arrival *x = (const arrival *)a, *y = (const arrival *)b;
    if (x->t < y->t) return -1;
    if (x->t > y->t) return 1;
    return (x->j > y->j) - (x->j < y->j);
}

void oracle_time_intensity(const oracle_geom *g, const double *atoms, const double *times, int64_t n, double c,
                           double tau, double *out, int nthreads)
{
    const int64_t npix = g->n0 * g->n1;
    if (nthreads <= 0) nthreads = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    {
        arrival *arr = (arrival *)malloc(sizeof(arrival) * (size_t)(n > 0 ? n : 1));
        double *are = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        double *aim = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
#pragma omp for schedule(dynamic, 4)
        for (int64_t p = 0; p < npix; ++p) {
            const int64_t x = p / g->n1, y = p % g->n1;
            double px = ((double)x - (double)g->n0 / 2.0) * g->pixelsize + g->off0;
            double py = ((double)y - (double)g->n1 / 2.0) * g->pixelsize + g->off1;
            double pz = g->detz;
            if (g->rot) {
                const double *M = g->M;
                const double rx = M[0] * px + M[1] * py + M[2] * pz;
                const double ry = M[3] * px + M[4] * py + M[5] * pz;
                const double rz = M[6] * px + M[7] * py + M[8] * pz;
                px = rx; py = ry; pz = rz;
            }
            const double R = sqrt(px * px + py * py + pz * pz);
            for (int64_t j = 0; j < n; ++j) {
                const double *e = atoms + 4 * j;
                double s, t, ph, sc = 1.0;
                if (g->nf) {
                    const double dx = e[0] - px, dy = e[1] - py, dz = e[2] - pz;
                    const double d = sqrt(dx * dx + dy * dy + dz * dz);
                    s = d - R + e[2];
                    t = s / c + times[j];
                    ph = s * g->k + e[3];
                    if (g->scale) sc = 1.0 / d;
                } else {
                    s = e[0] * (px / R) + e[1] * (py / R) + e[2] * (pz / R - 1.0);
                    t = -s / c + times[j];
                    ph = e[3] - s * g->k;
                    if (g->scale) sc = 1.0 / R;
                }
                arr[j].t = t;
                arr[j].j = j;
                are[j] = cos(ph) * sc;
                aim[j] = sin(ph) * sc;
            }
            qsort(arr, (size_t)n, sizeof(arrival), by_time);
            double xr = 0.0, xi = 0.0, acc = 0.0;
            for (int64_t i = 0; i < n; ++i) {
                const int64_t j = arr[i].j;
                if (i == 0) {
                    xr = are[j];
                    xi = aim[j];
                } else {
                    const double d = exp(-(arr[i].t - arr[i - 1].t) / tau);
                    xr = are[j] + d * xr;
                    xi = aim[j] + d * xi;
                }
                const double m2 = xr * xr + xi * xi;
                if (i + 1 < n)
                    acc += -tau / 2.0 * m2 * expm1(-2.0 * (arr[i + 1].t - arr[i].t) / tau);
                else
                    acc += tau / 2.0 * m2;
            }
            out[p] = acc;
        }
        free(arr); free(are); free(aim);
    }
}
