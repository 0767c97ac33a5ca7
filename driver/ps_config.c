/* ps_config.c -- ini keys -> ps_config, as confObj_new (confObj.c:27-110). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ps_driver.h"

#define GET(trg, fn, ini, key, sec)                                                       \
    do {                                                                                  \
        if (!fn(ini, key, sec, trg)) {                                                    \
            fprintf(stderr, "FATAL:  Could not get %s from section %s.\n", key, sec);     \
            exit(EXIT_FAILURE);                                                           \
        }                                                                                 \
    } while (0)

ps_config *ps_read_config(const char *path)
{
    ps_ini *ini = ps_ini_open(path);
    if (!ini) {
        fprintf(stderr, "FATAL:  Could not open %s for reading.\n", path);
        exit(EXIT_FAILURE);
    }
    ps_config *c = (ps_config *)calloc(1, sizeof(ps_config));
    char *kcase = NULL;
    GET(&c->grid_size, ps_ini_get_int32, ini, "gridsize", "Grid");
    GET(&c->box_size, ps_ini_get_double, ini, "boxsize", "Grid");
    GET(&c->gas_inputs_in_dp, ps_ini_get_int32, ini, "gasInputsInDoublePrecision", "Grid");
    GET(&c->ion_inputs_in_dp, ps_ini_get_int32, ini, "ionInputsInDoublePrecision", "Grid");
    GET(&c->density_file, ps_ini_get_string, ini, "densityFile", "Grid");
    GET(&c->ion_file, ps_ini_get_string, ini, "ionFile", "Grid");
    GET(&c->output_dir, ps_ini_get_string, ini, "output_dir", "Output");
    GET(&c->output_basename, ps_ini_get_string, ini, "output_basename", "Output");
    GET(&c->write_numpolygons, ps_ini_get_int32, ini, "write_numpolygons", "Output");
    GET(&c->h, ps_ini_get_double, ini, "hubble_h", "Cosmology");
    GET(&c->omega_b, ps_ini_get_double, ini, "omega_b", "Cosmology");
    GET(&c->omega_m, ps_ini_get_double, ini, "omega_m", "Cosmology");
    GET(&c->omega_l, ps_ini_get_double, ini, "omega_l", "Cosmology");
    GET(&c->sigma8, ps_ini_get_double, ini, "sigma8", "Cosmology");
    GET(&c->Y, ps_ini_get_double, ini, "Y", "Cosmology");
    GET(&c->which_field, ps_ini_get_string, ini, "whichField", "Polyspectrum");
    GET(&c->calc_phases_only, ps_ini_get_int32_ifavail, ini, "useOnlyPhase", "Polyspectrum");
    GET(&c->n, ps_ini_get_int32, ini, "n", "Polyspectrum");
    GET(&c->equilateral, ps_ini_get_int32, ini, "equilateral", "Polyspectrum");
    GET(&c->k1, ps_ini_get_double, ini, "k1", "Polyspectrum");
    GET(&c->k2, ps_ini_get_double, ini, "k2", "Polyspectrum");
    GET(&c->num_values, ps_ini_get_int32, ini, "numValues", "Polyspectrum");
    GET(&c->theta, ps_ini_get_double, ini, "theta", "Polyspectrum");
    GET(&c->kbinwidth, ps_ini_get_double, ini, "kbinwidth", "Polyspectrum");
    GET(&c->dcosTheta, ps_ini_get_double, ini, "cosThetaBinwidth", "Polyspectrum");
    GET(&kcase, ps_ini_get_string, ini, "kbinningCase", "Polyspectrum");
    if (strcmp(kcase, "DERIVED_BINNING") == 0) c->kbinningCase = 1;
    else if (strcmp(kcase, "GIVEN_BINNING") == 0) c->kbinningCase = 2;
    else if (strcmp(kcase, "WATKINSON") == 0) c->kbinningCase = 3;
    else c->kbinningCase = 0;  /* anything else is DEFAULT (confObj.c:98-105) */
    free(kcase);
    /* optional GPU keys */
    char *prec = NULL;
    c->precision = 1;
    c->batched = 1;
    if (ps_ini_get_string(ini, "precision", "GPU", &prec)) {
        if (strncmp(prec, "MIXED", 5) == 0) c->precision = 0;
        free(prec);
    }
    int32_t b;
    if (ps_ini_get_int32(ini, "batched", "GPU", &b)) c->batched = (b != 0);
    c->quirks = 1;
    if (ps_ini_get_int32(ini, "referenceQuirks", "GPU", &b)) c->quirks = (b != 0);
    ps_ini_close(ini);
    return c;
}

void ps_free_config(ps_config *c)
{
    if (!c) return;
    free(c->density_file);
    free(c->ion_file);
    free(c->output_dir);
    free(c->output_basename);
    free(c->which_field);
    free(c);
}
