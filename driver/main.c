/* main.c -- `polyspectrum iniFile.ini` (the reference's main.c:31-86). */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "ps_driver.h"

int main(int argc, char *argv[])
{
    double t1 = (double)time(NULL), t2;
    if (argc != 2) {
        printf("cifog: (C)  - Use at own risk...\n");
        printf("USAGE:\n");
        printf("cifog iniFile\n");
        exit(EXIT_FAILURE);
    }
    ps_config *c = ps_read_config(argv[1]);
    ps_run(c);
    ps_free_config(c);
    t2 = (double)time(NULL);
    printf("Execution took %f s\n", t2 - t1);
    return 0;
}
