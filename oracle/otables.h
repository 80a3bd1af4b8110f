/* otables.h -- table loader of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 * Reads woof_nam_amd/data/rrtmg_tables.bin (format: woof_nam_amd/data/make_tables.py) and
 * promotes every float32 table to double. */
#ifndef WOOFRAD_OTABLES_H
#define WOOFRAD_OTABLES_H

#include <stdint.h>

typedef struct {
    char name[48];
    int dtype;      /* 0 float32 -> stored as double in d; 1 int32 -> i */
    int ndim;
    int dims[4];
    long n;
    double *d;
    int32_t *i;
} ot_entry;

typedef struct {
    int n;
    ot_entry *e;
} ot_file;

int ot_load(const char *path, ot_file *f);
void ot_free(ot_file *f);
const double *ot_d(const ot_file *f, const char *name, long expect_n); /* NULL if absent */
const int32_t *ot_i(const ot_file *f, const char *name, long expect_n);

#endif
