/*
 * ps_driver.h -- the drop-in driver: `polyspectrum iniFile.ini`.
 *
 * Same command line, ini keys, stdout lines and output files as the
 * reference (main.c, run.c, confObj.c, kvectors.c, input_field.c, grid.c);
 * the estimator runs on the GPU through include/polyspectrum_gpu.h.
 */
#ifndef PS_DRIVER_H
#define PS_DRIVER_H

#include "ps_ini.h"

#define PS_MAXLENGTH 1024

/* confObj_struct (confObj.h:24-58), plus the optional GPU keys */
typedef struct {
    int grid_size;
    double box_size;
    int gas_inputs_in_dp;
    int ion_inputs_in_dp;
    char *density_file;
    char *ion_file;
    double h, omega_b, omega_m, omega_l, sigma8, Y; /* read, unused (confObj.c:62-73) */
    char *which_field;
    int calc_phases_only; /* -1 when useOnlyPhase is absent */
    int n;
    int equilateral;
    double k1, k2;
    int num_values;
    double theta;
    double kbinwidth;
    int kbinningCase;
    double dcosTheta;
    char *output_dir;
    char *output_basename;
    int write_numpolygons;
    /* optional [GPU] section (not in the reference): precision FP64|MIXED,
     * batched 0|1 (1: every configuration in one captured launch sequence),
     * referenceQuirks 1|0 (1, the default: the reference's behaviour where it
     * is a defect, e.g. n = 2 GIVEN/WATKINSON binning keeps the last leg's
     * kmin/kmax at k[1], kvectors.c:92-97 with run.c:56-83; 0: corrected) */
    int precision;
    int batched;
    int quirks;
} ps_config;

/* readConfObj (confObj.c:131-150): exits with the reference's FATAL messages */
ps_config *ps_read_config(const char *path);
void ps_free_config(ps_config *c);

/* kvectors_t (kvectors.h:6-19) */
typedef struct {
    int n;
    double kpolygon[3];
    double kpolygonMin[3];
    double kpolygonMax[3];
    double kbinwidth;
    int kbinningCase;
    double dcosTheta;
    int numValues;
    double *theta;
    double *k;
} ps_kvectors;

ps_kvectors *ps_read_kvectors(const ps_config *c);
void ps_free_kvectors(ps_kvectors *kv);

/* run (run.c:29-125) */
int ps_run(const ps_config *c);

#endif
