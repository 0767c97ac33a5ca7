/*
 * ps_run.c -- k-lists, input grids, the per-configuration loop and the
 * output files of the reference driver (kvectors.c, input_field.c, grid.c,
 * run.c), with the estimator on the GPU (include/polyspectrum_gpu.h).
 */
#define _GNU_SOURCE
#include <dirent.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/polyspectrum_gpu.h"
#include "ps_driver.h"

#define PI acos(-1.0)

/* ------------------------------------------------------------------ k-lists */
static double calc_k3(double k1, double k2, double cos_theta)
{
    return sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * cos_theta); /* kvectors.c:226-231 */
}

static double *zeros(int n)
{
    double *p = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    if (!p) {
        fprintf(stderr, "polyspectrum: out of host memory\n");
        exit(EXIT_FAILURE);
    }
    return p;
}

ps_kvectors *ps_read_kvectors(const ps_config *c)
{
    ps_kvectors *kv = (ps_kvectors *)calloc(1, sizeof(ps_kvectors));
    const int n = c->n;
    int numValues = c->num_values;
    kv->n = n;
    kv->kbinwidth = c->kbinwidth;
    kv->kbinningCase = c->kbinningCase;
    kv->dcosTheta = c->dcosTheta;
    kv->numValues = numValues;
    if (n < 2) {
        fprintf(stderr, "No valid input for n being the number of k-vectors, as their sum is unequal to zero: n>=2\n");
        exit(EXIT_FAILURE);
    } else if (n == 2) {
        if (c->num_values > 1) {
            /* N/2 bins k_i = (i+1) k_F (kvectors.c:200-211) */
            kv->numValues = c->grid_size / 2;
            kv->k = zeros(kv->numValues > 2 ? kv->numValues : 2);
            for (int i = 0; i < c->grid_size / 2; i++) kv->k[i] = 2. * PI / c->box_size * (double)(i + 1);
        } else {
            if (kv->numValues != 1) {
                fprintf(stderr, "polyspectrum: numValues must be >= 1\n");
                exit(EXIT_FAILURE);
            }
            kv->k = zeros(2); /* the reference writes k[1] into a 1-element array */
            kv->k[0] = c->k1;
            kv->k[1] = c->k1;
        }
        for (int i = 0; i < n; i++) {
            kv->kpolygon[i] = kv->k[i];
            kv->kpolygonMin[i] = kv->k[i];
            kv->kpolygonMax[i] = kv->k[i];
        }
    } else if (n == 3) {
        kv->kpolygon[0] = c->k1;
        kv->kpolygon[1] = c->k2;
        if (c->equilateral == 1) {
            if (kv->numValues <= 0) {
                fprintf(stderr, "polyspectrum: numValues must be >= 1\n");
                exit(EXIT_FAILURE);
            }
            /* log-spaced k_F (N/2)^((i+1)/numValues) (kvectors.c:213-224) */
            kv->k = zeros(numValues);
            const double kL = 2. * PI / c->box_size;
            for (int i = 0; i < numValues; i++)
                kv->k[i] = kL * pow(c->grid_size / 2, (double)(i + 1) / (double)numValues);
            kv->theta = zeros(1);
            kv->theta[0] = PI / 3.;
        } else if (c->num_values > 1) {
            /* theta_i = pi i / numValues (kvectors.c:172-198) */
            kv->theta = zeros(numValues);
            kv->k = zeros(numValues);
            for (int i = 0; i < numValues; i++) {
                kv->theta[i] = PI * (double)i / (double)numValues;
                kv->k[i] = calc_k3(c->k1, c->k2, cos(kv->theta[i]));
            }
        } else {
            if (kv->numValues != 1) {
                fprintf(stderr, "polyspectrum: numValues must be >= 1\n");
                exit(EXIT_FAILURE);
            }
            kv->theta = zeros(1);
            kv->k = zeros(1);
            kv->theta[0] = c->theta;
            kv->k[0] = calc_k3(c->k1, c->k2, cos(c->theta));
        }
    } else {
        printf("Not implemented yet\n");
        exit(EXIT_FAILURE);
    }
    return kv;
}

void ps_free_kvectors(ps_kvectors *kv)
{
    if (!kv) return;
    free(kv->theta);
    free(kv->k);
    free(kv);
}

/* kvectors.c:233-247 */
static double calc_k3min(const ps_kvectors *kv, int i)
{
    double ct = fmin(1., cos(kv->theta[i]) + 0.5 * kv->dcosTheta);
    return calc_k3(kv->kpolygon[0], kv->kpolygon[1], ct);
}
static double calc_k3max(const ps_kvectors *kv, int i)
{
    double ct = fmax(-1., cos(kv->theta[i]) - 0.5 * kv->dcosTheta);
    return calc_k3(kv->kpolygon[0], kv->kpolygon[1], ct);
}

/* ------------------------------------------------------------------ input */
/* read_grid / read_grid_doubleprecision (grid.c:121-228): raw N^3, no header */
static double *read_grid(const char *file, int N, int dp)
{
    const size_t cnt = (size_t)N * N * N;
    double *out = (double *)malloc(cnt * sizeof(double));
    FILE *f = fopen(file, "rb");
    if (!out || !f) {
        fprintf(stderr, "polyspectrum: cannot read input grid %s\n", file);
        exit(EXIT_FAILURE);
    }
    size_t got;
    if (dp == 1) {
        got = fread(out, sizeof(double), cnt, f);
    } else {
        float *tmp = (float *)malloc(cnt * sizeof(float));
        got = fread(tmp, sizeof(float), cnt, f);
        for (size_t i = 0; i < cnt; i++) out[i] = (double)tmp[i];
        free(tmp);
    }
    fclose(f);
    if (got != cnt) {
        fprintf(stderr, "polyspectrum: %s holds %zu of %zu grid values\n", file, got, cnt);
        exit(EXIT_FAILURE);
    }
    return out;
}

/* get_FT_field (input_field.c:40-69): the selected real field */
static double *input_field(const ps_config *c)
{
    printf("Option: %s\t", c->which_field);
    const int N = c->grid_size;
    if (strcmp(c->which_field, "DENS") == 0) return read_grid(c->density_file, N, c->gas_inputs_in_dp);
    if (strcmp(c->which_field, "XHII") == 0) return read_grid(c->ion_file, N, c->ion_inputs_in_dp);
    if (strcmp(c->which_field, "XHI_DENS") == 0) {
        double *rho = read_grid(c->density_file, N, c->gas_inputs_in_dp);
        double *x = read_grid(c->ion_file, N, c->ion_inputs_in_dp);
        const size_t cnt = (size_t)N * N * N;
        for (size_t i = 0; i < cnt; i++) rho[i] = rho[i] * (1. - x[i]); /* input_field.c:33 */
        free(x);
        return rho;
    }
    fflush(stdout);
    fprintf(stderr, "Not supported option. Please use 'DENS', 'XHII' or 'XHI_DENS'\n");
    exit(EXIT_FAILURE);
}

/* ------------------------------------------------------------------ output */
static void file_ending(const ps_config *c, char *ending)
{
    ending[0] = '\0';
    if (c->n == 2) sprintf(ending, "_ps.dat");
    else if (c->n == 3) {
        if (c->equilateral == 1) sprintf(ending, "_bs_equilateral.dat");
        else if (c->num_values == 1)
            sprintf(ending, "_bs_k_%4.2e_%4.2e_theta_%4.2e.dat", c->k1, c->k2, c->theta);
        else sprintf(ending, "_bs_k_%4.2e_%4.2e.dat", c->k1, c->k2);
    }
}

/* save_polyspectrum / save_polyspectrum_numpolygons (run.c:128-222) */
static void save(const ps_config *c, int num, const double *theta, const double *k, const double *val,
                 const double *npoly)
{
    char ending[PS_MAXLENGTH];
    file_ending(c, ending);
    size_t len = strlen(c->output_dir) + strlen(c->output_basename) + strlen(ending) + 2;
    char *fname = (char *)malloc(len);
    snprintf(fname, len, "%s/%s%s", c->output_dir, c->output_basename, ending);
    printf("FILENAME : %s\n", fname);
    FILE *f = fopen(fname, "wb");
    if (!f) {
        fprintf(stderr, "polyspectrum: cannot write %s\n", fname);
        free(fname);
        exit(EXIT_FAILURE);
    }
    if (c->n == 2 || c->equilateral == 1) {
        fprintf(f, "# k [h^-1 Mpc]\t Polyspectrum_%d(k)\n", c->n);
        for (int i = 0; i < num; i++) {
            if (npoly) fprintf(f, "%e\t%e\t%e\n", k[i], val[i], npoly[i]);
            else fprintf(f, "%e\t%e\n", k[i], val[i]);
        }
    } else if (c->n == 3) {
        fprintf(f, "# theta [rad]\t k [h^-1 Mpc]\t Polyspectrum_%d(k)\n", c->n);
        for (int i = 0; i < num; i++) {
            if (npoly) fprintf(f, "%e\t%e\t%e\t%e\n", theta[i], k[i], val[i], npoly[i]);
            else fprintf(f, "%e\t%e\t%e\n", theta[i], k[i], val[i]);
        }
    }
    fclose(f);
    free(fname);
}

static int dir_exists(const char *d)
{
    DIR *p = opendir(d);
    if (p) {
        closedir(p);
        return 1;
    }
    return 0;
}

#define GPUCK(call)                                                                      \
    do {                                                                                 \
        int s_ = (call);                                                                 \
        if (s_ != PS_OK) {                                                               \
            fprintf(stderr, "polyspectrum: %s: %s\n", #call, ps_strerror(s_));           \
            exit(EXIT_FAILURE);                                                          \
        }                                                                                \
    } while (0)

/* ------------------------------------------------------------------ run */
int ps_run(const ps_config *c)
{
    ps_kvectors *kv = ps_read_kvectors(c);
    printf("kbinningCase = %d\n", kv->kbinningCase);
    if (!dir_exists(c->output_dir)) {
        fprintf(stderr, "Directory for output does not exist!\n");
        exit(EXIT_FAILURE);
    }
    const int N = c->grid_size;
    if (!ps_grid_supported(N)) {
        fprintf(stderr, "polyspectrum: gridsize %d has no GPU FFT plan (2^a*3^b, 16..512)\n", N);
        exit(EXIT_FAILURE);
    }
    double *field = input_field(c);
    printf(" Mean value from FT is %e\nComputing FFT... ", field[0]); /* input_field.c:71 prints cell 0 */
    fflush(stdout);
    ps_ctx *ctx = NULL;
    GPUCK(ps_ctx_create(N, c->box_size, c->precision ? PS_PREC_FP64 : PS_PREC_MIXED, 1, &ctx));
    GPUCK(ps_load_field(ctx, 0, field, 1, 0));
    printf("done.\n");
    free(field);
    if (c->calc_phases_only == 1) GPUCK(ps_set_phase_only(ctx, 1));

    const int num = kv->numValues;
    const int n = kv->n;
    double *val = zeros(num), *npoly = zeros(num), *sumD = zeros(num), *sumI = zeros(num);
    ps_polygon *polys = (ps_polygon *)calloc((size_t)num, sizeof(ps_polygon));
    /* the kpolygon updates of run.c:56-83, one snapshot per configuration */
    for (int i = 0; i < num; i++) {
        kv->kpolygon[n - 1] = kv->k[i];
        if (n == 2) kv->kpolygon[0] = kv->k[i];
        if (n == 2 && !c->quirks) {
            /* corrected: the last leg's GIVEN/WATKINSON range follows k[i]; the
             * reference keeps kpolygonMin/Max[1] = k[1] from kvectors.c:92-97 */
            kv->kpolygonMin[n - 1] = kv->k[i];
            kv->kpolygonMax[n - 1] = kv->k[i];
        }
        if (n == 3) {
            if (c->equilateral == 1) {
                for (int j = 0; j < n; j++) kv->kpolygon[j] = kv->k[i];
                if (kv->kbinningCase >= 2) {
                    kv->kpolygonMin[n - 1] = calc_k3min(kv, 0);
                    kv->kpolygonMax[n - 1] = calc_k3max(kv, 0);
                }
            } else if (kv->kbinningCase >= 2) {
                kv->kpolygonMin[n - 1] = calc_k3min(kv, i);
                kv->kpolygonMax[n - 1] = calc_k3max(kv, i);
            }
        }
        polys[i].n = n;
        polys[i].slot = 0;
        memcpy(polys[i].k, kv->kpolygon, sizeof(polys[i].k));
        memcpy(polys[i].kmin, kv->kpolygonMin, sizeof(polys[i].kmin));
        memcpy(polys[i].kmax, kv->kpolygonMax, sizeof(polys[i].kmax));
        if (!c->batched) {
            /* polyspectrum() writes kmin/kmax back for DERIVED binning */
            double s[2];
            GPUCK(ps_polyspectrum(ctx, 0, n, polys[i].k, kv->kpolygonMin, kv->kpolygonMax, c->kbinwidth,
                                  kv->kbinningCase, &val[i], s));
            sumD[i] = s[0];
            sumI[i] = s[1];
            if (c->write_numpolygons == 1)
                GPUCK(ps_num_polygons(ctx, n, polys[i].k, kv->kpolygonMin, kv->kpolygonMax, c->kbinwidth,
                                      kv->kbinningCase, &npoly[i]));
        }
    }
    if (c->batched) {
        ps_batch *b = NULL;
        double *sums = zeros(2 * num);
        GPUCK(ps_batch_create(ctx, polys, num, c->kbinwidth, kv->kbinningCase, &b));
        GPUCK(ps_batch_execute(b, val, npoly, sums));
        for (int i = 0; i < num; i++) {
            sumD[i] = sums[2 * i];
            sumI[i] = sums[2 * i + 1];
        }
        ps_batch_destroy(b);
        free(sums);
    }
    for (int i = 0; i < num; i++) {
        /* polyspectrum.c:142: Im/Re of the sums; the GPU path sums real fields */
        printf("rank %d: cimag(sumNumPolygons)/creal(sumNumPolygons) = %e\n\tcimag(sumValuesPolygons)/creal(sumValuesPolygons) = %e\n",
               0, 0.0 / sumI[i], 0.0 / sumD[i]);
        if (n == 2) printf("k = %e\t P(k) = %e\t", kv->k[i], val[i]);
        if (n == 3)
            printf("k1 = %e\t k2 = %e\tk3 = %e\t B(k) = %e\t", polys[i].k[0], polys[i].k[1], kv->k[i], val[i]);
        if (c->write_numpolygons == 1) printf("Npolygons = %e\n", npoly[i]);
        printf("\n");
    }
    save(c, num, kv->theta, kv->k, val, c->write_numpolygons == 1 ? npoly : NULL);
    free(polys);
    free(val);
    free(npoly);
    free(sumD);
    free(sumI);
    ps_ctx_destroy(ctx);
    ps_free_kvectors(kv);
    return 0;
}
