/*
 * Plain-C restatement of Acoular's frequency-domain beamforming loop nests.
 *
 * TEST INFRASTRUCTURE ONLY: linked by the tests (as a second oracle, checked against the
 * numpy restatement) and timed by the benchmark's CPU-baseline leg.  The product path never
 * calls it.
 *
 * The loops follow the reference one to one:
 *   cref_beamformer_full_csm  <- _freqBeamformer_FullCSM   acoular/fastFuncs.py:238-311
 *       (float64 arithmetic, phase argument rounded to float32 and cos/sin taken in float32
 *        as at :259-261, Hermitian upper-triangle quadratic form :277-291, normalisations
 *        :293-310; OpenMP parallel over grid points like the numba prange at :255)
 *   cref_calc_csm             <- calcCSM                   acoular/fastFuncs.py:30-58
 *       (prange over frequencies, upper triangle incl. diagonal)
 *   cref_dist_mat             <- dist_mat                  acoular/environments.py:65-103
 *   cref_psf                  <- calcPointSpreadFunction   acoular/fastFuncs.py:563-801
 *       (float32-rounded phase difference and float32 cos/sin, float64 sums; OpenMP over grid
 *        points like the guvectorize target)
 *   cref_damas_gauss_seidel   <- damasSolverGaussSeidel    acoular/fastFuncs.py:820-855
 *       (the same row order and summation order; OpenMP over frequencies)
 *
 * Complex arrays are interleaved (re, im) doubles, C-contiguous, as numpy complex128.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

int cref_num_threads(void)
{
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* threads of the following parallel loops (OpenMP); no effect without OpenMP */
void cref_set_threads(int n)
{
#ifdef _OPENMP
    extern void omp_set_num_threads(int);
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

void cref_dist_mat(const double *gpos, const double *mpos, int64_t ng, int64_t nm, double *rm)
{
    const double *m0 = mpos, *m1 = mpos + nm, *m2 = mpos + 2 * nm;
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < ng; ++n) {
        double g0 = gpos[n], g1 = gpos[ng + n], g2 = gpos[2 * ng + n];
        for (int64_t m = 0; m < nm; ++m) {
            double d0 = g0 - m0[m], d1 = g1 - m1[m], d2 = g2 - m2[m];
            rm[n * nm + m] = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        }
    }
}

void cref_beamformer_full_csm(const double *csm, const double *r0, const double *rm, double k,
                              double norm, int r_diag, int steer_type, int64_t ng, int64_t nm,
                              double *result, double *normalize_steer)
{
    const int st2 = steer_type == 2;
    const int st34 = steer_type == 3 || steer_type == 4;
#pragma omp parallel
    {
        double *sv = (double *)malloc(sizeof(double) * 2 * (size_t)nm);
#pragma omp for schedule(static)
        for (int64_t gi = 0; gi < ng; ++gi) {
            const double *r = rm + gi * nm;
            double help = 0.0;
            for (int64_t m = 0; m < nm; ++m) {
                float arg = (float)(k * r[m]);
                sv[2 * m] = (double)cosf(arg);
                sv[2 * m + 1] = -(double)sinf(arg);
            }
            if (st2) {
                for (int64_t m = 0; m < nm; ++m) {
                    help += r[m] * r[m];
                    sv[2 * m] *= r[m];
                    sv[2 * m + 1] *= r[m];
                }
            }
            if (st34) {
                for (int64_t m = 0; m < nm; ++m) {
                    help += 1.0 / (r[m] * r[m]);
                    sv[2 * m] /= r[m];
                    sv[2 * m + 1] /= r[m];
                }
            }
            double q = 0.0;
            for (int64_t j = 0; j < nm; ++j) {
                double lr = 0.0, li = 0.0; /* sum_{i<j} C_ij conj(v_i) */
                for (int64_t i = 0; i < j; ++i) {
                    double cr = csm[2 * (i * nm + j)], ci = csm[2 * (i * nm + j) + 1];
                    double vr = sv[2 * i], vi = -sv[2 * i + 1];
                    lr += cr * vr - ci * vi;
                    li += cr * vi + ci * vr;
                }
                q += 2.0 * (lr * sv[2 * j] - li * sv[2 * j + 1]);
            }
            if (!r_diag) {
                for (int64_t m = 0; m < nm; ++m) {
                    double cr = csm[2 * (m * nm + m)];
                    q += cr * (sv[2 * m] * sv[2 * m] + sv[2 * m + 1] * sv[2 * m + 1]);
                }
            }
            double nf;
            switch (steer_type) {
            case 1:
                nf = (double)nm;
                normalize_steer[gi] = 1.0 / nm;
                result[gi] = q / (nf * nf) * norm;
                break;
            case 2:
                nf = nm * r0[gi];
                normalize_steer[gi] = help / (nf * nf);
                result[gi] = q / (nf * nf) * norm;
                break;
            case 3:
                nf = r0[gi] * help;
                normalize_steer[gi] = 1.0 / (r0[gi] * r0[gi]) / help;
                result[gi] = q / (nf * nf) * norm;
                break;
            default:
                nf = nm * help;
                normalize_steer[gi] = 1.0 / nm;
                result[gi] = q / nf * norm;
                break;
            }
        }
        free(sv);
    }
}

void cref_calc_csm(double *csm, const double *spec, int64_t nf, int64_t nm)
{
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < nf; ++f) {
        double *c = csm + 2 * f * nm * nm;
        const double *s = spec + 2 * f * nm;
        for (int64_t col = 0; col < nm; ++col) {
            double tr = s[2 * col], ti = -s[2 * col + 1];
            for (int64_t row = 0; row <= col; ++row) {
                double sr = s[2 * row], si = s[2 * row + 1];
                c[2 * (row * nm + col)] += tr * sr - ti * si;
                c[2 * (row * nm + col) + 1] += tr * si + ti * sr;
            }
        }
    }
}

/* psf [ng][ns]: column c is the PSF of source grid point ind[c] */
void cref_psf(const double *r0, const double *rm, int64_t ng, int64_t nm, double k, const int64_t *ind, int64_t ns,
              int steer_type, double *psf)
{
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < ng; ++g) {
        const double *rg = rm + g * nm;
        double help = 0.0;
        for (int64_t m = 0; m < nm; ++m) help += 1.0 / (rg[m] * rg[m]);
        for (int64_t c = 0; c < ns; ++c) {
            const int64_t s = ind[c];
            const double *rs = rm + s * nm;
            double pr = 0.0, pi = 0.0;
            for (int64_t m = 0; m < nm; ++m) {
                float arg = (float)(k * (rg[m] - rs[m]));
                double w = 1.0 / rs[m];
                if (steer_type == 2) w *= rg[m];
                if (steer_type == 3 || steer_type == 4) w /= rg[m];
                pr += (double)cosf(arg) * w;
                pi -= (double)sinf(arg) * w;
            }
            const double sq = pr * pr + pi * pi;
            double nf;
            switch (steer_type) {
            case 1:
                nf = r0[s] / nm;
                psf[g * ns + c] = sq * nf * nf;
                break;
            case 2:
                nf = r0[s] / r0[g] / nm;
                psf[g * ns + c] = sq * nf * nf;
                break;
            case 3:
                nf = r0[s] / r0[g] / help;
                psf[g * ns + c] = sq * nf * nf;
                break;
            default:
                psf[g * ns + c] = sq * (r0[s] * r0[s]) / nm / help;
                break;
            }
        }
    }
}

/* A [nf][ng][ng], dirty [nf][ng], x [nf][ng] (start value in, solution out) */
void cref_damas_gauss_seidel(const double *A, const double *dirty, double *x, int64_t nf, int64_t ng,
                             int64_t n_iter, double relax)
{
#pragma omp parallel for schedule(dynamic)
    for (int64_t f = 0; f < nf; ++f) {
        const double *a = A + f * ng * ng;
        const double *y = dirty + f * ng;
        double *xs = x + f * ng;
        for (int64_t it = 0; it < n_iter; ++it) {
            for (int64_t i = 0; i < ng; ++i) {
                double s = 0.0;
                for (int64_t j = 0; j < ng; ++j) s += a[i * ng + j] * xs[j];
                s -= a[i * ng + i] * xs[i];
                s = (1 - relax) * xs[i] + relax * (y[i] - s);
                xs[i] = s > 0.0 ? s : 0.0;
            }
        }
    }
}
