/* otables.c -- table loader of the CPU oracle (TEST INFRASTRUCTURE ONLY). */
#include "otables.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int ot_load(const char *path, ot_file *f)
{
    memset(f, 0, sizeof(*f));
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    fseek(fp, 0, SEEK_END);
    long sz = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    unsigned char *buf = (unsigned char *)malloc((size_t)sz);
    if (!buf || fread(buf, 1, (size_t)sz, fp) != (size_t)sz) {
        fclose(fp);
        free(buf);
        return -1;
    }
    fclose(fp);
    if (sz < 12 || memcmp(buf, "WRTBL001", 8) != 0) {
        free(buf);
        return -2;
    }
    uint32_t n;
    memcpy(&n, buf + 8, 4);
    f->n = (int)n;
    f->e = (ot_entry *)calloc(n, sizeof(ot_entry));
    const unsigned char *p = buf + 12;
    for (uint32_t k = 0; k < n; ++k, p += 48 + 4 + 4 + 16 + 8 + 8) {
        ot_entry *e = &f->e[k];
        memcpy(e->name, p, 48);
        e->name[47] = 0;
        uint32_t dt, nd, dims[4];
        uint64_t off, cnt;
        memcpy(&dt, p + 48, 4);
        memcpy(&nd, p + 52, 4);
        memcpy(dims, p + 56, 16);
        memcpy(&off, p + 72, 8);
        memcpy(&cnt, p + 80, 8);
        e->dtype = (int)dt;
        e->ndim = (int)nd;
        for (int j = 0; j < 4; ++j) e->dims[j] = (int)dims[j];
        e->n = (long)cnt;
        if (off + cnt * 4 > (uint64_t)sz) {
            free(buf);
            return -3;
        }
        if (dt == 0) {
            e->d = (double *)malloc(sizeof(double) * cnt);
            const float *src = (const float *)(buf + off);
            for (uint64_t j = 0; j < cnt; ++j) e->d[j] = (double)src[j];
        } else {
            e->i = (int32_t *)malloc(sizeof(int32_t) * cnt);
            memcpy(e->i, buf + off, 4 * cnt);
        }
    }
    free(buf);
    return 0;
}

void ot_free(ot_file *f)
{
    for (int k = 0; k < f->n; ++k) {
        free(f->e[k].d);
        free(f->e[k].i);
    }
    free(f->e);
    memset(f, 0, sizeof(*f));
}

static const ot_entry *find(const ot_file *f, const char *name)
{
    for (int k = 0; k < f->n; ++k)
        if (strcmp(f->e[k].name, name) == 0) return &f->e[k];
    return NULL;
}

const double *ot_d(const ot_file *f, const char *name, long expect_n)
{
    const ot_entry *e = find(f, name);
    if (!e || e->dtype != 0 || (expect_n > 0 && e->n != expect_n)) return NULL;
    return e->d;
}

const int32_t *ot_i(const ot_file *f, const char *name, long expect_n)
{
    const ot_entry *e = find(f, name);
    if (!e || e->dtype != 1 || (expect_n > 0 && e->n != expect_n)) return NULL;
    return e->i;
}
