/*
 * Driver for the reference's vendored MTGP32 C library, built from the
 * reference's own source files where they lie (see oracle/Makefile) into
 * oracle/_ref/.  It is test infrastructure only: it writes the MTGP32-11213
 * parameter table as data and known-answer words that pin the oracle's
 * restatement and the GPU generator.
 *
 * Reference interfaces driven here:
 *   mtgp32_init_state     src/partfilt/opencl/mtgp/mtgp32-fast.c:119-137
 *   mtgp32_init           src/partfilt/opencl/mtgp/mtgp32-fast.c:94-104
 *   mtgp32_genrand_uint32 src/partfilt/opencl/mtgp/mtgp32-fast.h:252-267
 *   mtgp32_genrand_close1_open2  mtgp32-fast.h:278-293
 *   table mtgp32dc_params_fast_11213[200]  mtgp32dc-param-11213.c
 *
 * Usage:
 *   mtgp_ref params                       -> text table on stdout
 *   mtgp_ref init  SEED SET               -> 351 state words (hex) for set SET, seed SEED
 *   mtgp_ref u32   SEED SET COUNT         -> COUNT tempered uint32 words (hex)
 *   mtgp_ref f32   SEED SET COUNT         -> COUNT close1_open2 float bit patterns (hex)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include "mtgp32-fast.h"

extern mtgp32_params_fast_t mtgp32dc_params_fast_11213[];
extern const int mtgpdc_params_11213_num;

static void usage(void) {
    fprintf(stderr, "usage: mtgp_ref params | init SEED SET | u32 SEED SET COUNT | f32 SEED SET COUNT\n");
    exit(2);
}

int main(int argc, char** argv) {
    if (argc < 2) usage();
    if (strcmp(argv[1], "params") == 0) {
        printf("# MTGP32-11213 parameter sets: pos sh1 sh2 mask tbl[16] tmp_tbl[16] flt_tmp_tbl[16]\n");
        for (int i = 0; i < mtgpdc_params_11213_num; ++i) {
            const mtgp32_params_fast_t* p = &mtgp32dc_params_fast_11213[i];
            if (p->mexp != 11213) { fprintf(stderr, "bad mexp at %d\n", i); return 1; }
            printf("%d %d %d %08x", p->pos, p->sh1, p->sh2, p->mask);
            for (int j = 0; j < 16; ++j) printf(" %08x", p->tbl[j]);
            for (int j = 0; j < 16; ++j) printf(" %08x", p->tmp_tbl[j]);
            for (int j = 0; j < 16; ++j) printf(" %08x", p->flt_tmp_tbl[j]);
            printf("\n");
        }
        return 0;
    }
    if (argc < 4) usage();
    uint32_t seed = (uint32_t)strtoul(argv[2], NULL, 0);
    int set = atoi(argv[3]);
    if (set < 0 || set >= mtgpdc_params_11213_num) usage();
    const mtgp32_params_fast_t* p = &mtgp32dc_params_fast_11213[set];
    if (strcmp(argv[1], "init") == 0) {
        uint32_t st[351];
        mtgp32_init_state(st, p, seed);
        for (int i = 0; i < 351; ++i) printf("%08x\n", st[i]);
        return 0;
    }
    if (argc < 5) usage();
    int count = atoi(argv[4]);
    mtgp32_fast_t g;
    if (mtgp32_init(&g, p, seed) != 0) return 1;
    if (strcmp(argv[1], "u32") == 0) {
        for (int i = 0; i < count; ++i) printf("%08x\n", mtgp32_genrand_uint32(&g));
    } else if (strcmp(argv[1], "f32") == 0) {
        for (int i = 0; i < count; ++i) {
            float f = mtgp32_genrand_close1_open2(&g);
            uint32_t u;
            memcpy(&u, &f, 4);
            printf("%08x\n", u);
        }
    } else {
        usage();
    }
    mtgp32_free(&g);
    return 0;
}
