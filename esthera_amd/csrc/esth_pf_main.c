/*
 * esthera_pf - the reference's particle_filter command line on the engine.
 *
 * Flags as src/partfilt/opencl/pf_generic.c:1071-1138:
 *   -m a[:b]   particles per sub-filter, a..b doubling       (default 256)
 *   -N a[:b]   sub-filters, a..b doubling                    (default 16)
 *   -t a[:b]   particles exchanged per neighbour, a..b by 1  (default 1)
 *   -r x[:y]   resampling threshold, x..y by 0.1 (0.01 with -A; default 1.0,
 *              0.0 with -A as the ADAPTIVE_RESAMPLING build, :1084-1091)
 *   -l n       runs of each configuration                    (default 1)
 *   -d p:d     device d (one platform: p must be 0)
 *   -c         a CPU device: not available, exits with an error
 *   -v         print the device and memory budget
 * and the switches the reference selects at compile time:
 *   -M model   robotarm5 (default) | robotarm5_camera | unicycle10
 *   -R kind    multinomial (default) | systematic | alias
 *   -A         adaptive resampling (alias; resample iff sum p^2 > threshold)
 *   -s seed    MTGP seed (default 1234)
 *   -T         per-step lines instead of per-configuration ones (DEBUG_TRACE)
 * Output: "N m r t e" then one line per run (:814, the mean estimate error of
 * the best particle), or with -T "ex ey ew ee ax ay" then one line per step
 * (print_estimate, cuda/pf_model_soa_packed_code.cuh:237-240).
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "esthera.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static void usage(const char* prog) {
    fprintf(stderr,
            "Usage: %s [-m #particles] [-N #cores] [-t #transfer] [-r resampling_threshold] [-l loop_count] "
            "[-d platform_id:device_id] [-c] [-v] [-M model] [-R resampler] [-A] [-s seed] [-T] input_file\n",
            prog);
}

static bool parse_range(const char* s, int* a, int* b) {
    char* end = NULL;
    *a = (int)strtol(s, &end, 10);
    if (end == s) return false;
    if (*end == ':') {
        const char* r = end + 1;
        *b = (int)strtol(r, &end, 10);
        if (end == r) return false;
    } else {
        *b = *a;
    }
    return *end == '\0';
}

static bool parse_frange(const char* s, double* a, double* b) {
    char* end = NULL;
    *a = strtod(s, &end);
    if (end == s) return false;
    if (*end == ':') {
        const char* r = end + 1;
        *b = strtod(r, &end);
        if (end == r) return false;
    } else {
        *b = *a;
    }
    return *end == '\0';
}

struct trace {
    int model;
    int rows;
    esth_robotarm5_measurement* ra_meas;
    esth_robotarm5_control* ra_ctrl;
    esth_unicycle10_measurement* u_meas;
    esth_unicycle_control* u_ctrl;
    float *dt, *truth;   /* 9 (arm) or 3 (unicycle) floats per row */
};

static int count_rows(const char* path) {
    FILE* fh = fopen(path, "r");
    if (!fh) return -1;
    int n = 0, c, last = '\n';
    while ((c = fgetc(fh)) != EOF) {
        if (c == '\n') ++n;
        last = c;
    }
    if (last != '\n') ++n;
    fclose(fh);
    return n;
}

static int load_trace(const char* path, struct trace* tr) {
    const int cap = count_rows(path);
    if (cap < 0) {
        fprintf(stderr, "could not open: %s\n", path);
        return ESTH_ERR_IO;
    }
    const size_t n = cap > 0 ? (size_t)cap : 1;
    tr->dt = calloc(n, sizeof(float));
    if (tr->model == ESTH_MODEL_UNICYCLE10) {
        tr->u_meas = calloc(n, sizeof(*tr->u_meas));
        tr->u_ctrl = calloc(n, sizeof(*tr->u_ctrl));
        tr->truth = calloc(n * 3, sizeof(float));
        if (!tr->dt || !tr->u_meas || !tr->u_ctrl || !tr->truth) return ESTH_ERR_NOMEM;
        return esth_trace_load_unicycle10(path, cap, tr->u_meas, tr->u_ctrl, tr->dt, tr->truth, &tr->rows);
    }
    tr->ra_meas = calloc(n, sizeof(*tr->ra_meas));
    tr->ra_ctrl = calloc(n, sizeof(*tr->ra_ctrl));
    tr->truth = calloc(n * 9, sizeof(float));
    if (!tr->dt || !tr->ra_meas || !tr->ra_ctrl || !tr->truth) return ESTH_ERR_NOMEM;
    return esth_trace_load_robotarm(path, cap, tr->ra_meas, tr->ra_ctrl, tr->dt, tr->truth, &tr->rows);
}

static void free_trace(struct trace* tr) {
    free(tr->ra_meas);
    free(tr->ra_ctrl);
    free(tr->u_meas);
    free(tr->u_ctrl);
    free(tr->dt);
    free(tr->truth);
}

/* one run of one configuration; returns the mean error over the trace */
static int run(esth_engine* e, const esth_filter_desc* d, const void* params, size_t psize, const struct trace* tr,
               bool per_step, double* mean_error) {
    esth_filter* f = NULL;
    int rc = esth_filter_create(e, d, params, psize, &f);
    if (rc != ESTH_OK) return rc;
    if (tr->model == ESTH_MODEL_UNICYCLE10) {
        const float init[3] = {20.0f, 40.0f, (float)(-M_PI / 2)};   /* particle_filter.cu:712-721 */
        rc = esth_filter_init_uniform(f, init);
    } else {
        const float a = (float)(0.75 * M_PI);   /* pf_generic.c:346-361 */
        const float init[9] = {a, a, a, a, a, 7.0f, 0.0f, 0.0f, 0.0f};
        rc = esth_filter_init_uniform(f, init);
    }
    double err_sum = 0.0;
    for (int i = 0; rc == ESTH_OK && i < tr->rows; ++i) {
        if (tr->model == ESTH_MODEL_UNICYCLE10)
            rc = esth_filter_step(f, &tr->u_ctrl[i], &tr->u_meas[i], tr->dt[i]);
        else
            rc = esth_filter_step(f, &tr->ra_ctrl[i], &tr->ra_meas[i], tr->dt[i]);
        esth_estimate est;
        if (rc == ESTH_OK) rc = esth_filter_estimate(f, &est);
        if (rc != ESTH_OK) break;
        double err;
        const float* act;
        if (tr->model == ESTH_MODEL_UNICYCLE10) {
            act = &tr->truth[(size_t)i * 3];
            const double ex = (double)est.best[0] - act[0], ey = (double)est.best[1] - act[1];
            err = sqrt(ex * ex + ey * ey);
        } else {
            act = &tr->truth[(size_t)i * 9];
            float best[9];
            for (int c = 0; c < 9; ++c) best[c] = est.best[c];
            best[7] = 0.0f;   /* the reference reads back x, y and the angles only (pf_generic.c:649-663) */
            best[8] = 0.0f;
            err = esth_robotarm5_estimate_error(best, act);
        }
        err_sum += err;
        if (per_step) {
            const int ix = tr->model == ESTH_MODEL_UNICYCLE10 ? 0 : 5;
            printf("%f %f %f %f %f %f\n", est.best[ix], est.best[ix + 1], est.best_weight, err, act[ix], act[ix + 1]);
        }
    }
    esth_filter_destroy(f);
    if (rc == ESTH_OK) *mean_error = tr->rows ? err_sum / tr->rows : 0.0;
    return rc;
}

int main(int argc, char* argv[]) {
    int t0 = 1, t1 = 1, g0 = 16, g1 = 16, m0 = 256, m1 = 256, loops = 1, device = 0;
    double r0 = 1.0, r1 = 1.0, rinc = 0.1;
    bool r_given = false, verbose = false, per_step = false, adaptive = false;
    int resampler = ESTH_RESAMPLE_MULTINOMIAL;
    uint32_t seed = 1234;
    struct trace tr;
    memset(&tr, 0, sizeof(tr));
    tr.model = ESTH_MODEL_ROBOTARM5;
    int opt;
    while ((opt = getopt(argc, argv, "m:N:t:r:l:d:cvM:R:As:Th")) != -1) {
        switch (opt) {
        case 'm': if (!parse_range(optarg, &m0, &m1)) { usage(argv[0]); return EXIT_FAILURE; } break;
        case 'N': if (!parse_range(optarg, &g0, &g1)) { usage(argv[0]); return EXIT_FAILURE; } break;
        case 't': if (!parse_range(optarg, &t0, &t1)) { usage(argv[0]); return EXIT_FAILURE; } break;
        case 'r':
            if (!parse_frange(optarg, &r0, &r1)) { usage(argv[0]); return EXIT_FAILURE; }
            r_given = true;
            break;
        case 'l': loops = atoi(optarg); break;
        case 'd': {
            int platform = 0;
            if (sscanf(optarg, "%d:%d", &platform, &device) != 2 || platform != 0 || device < 0) {
                fprintf(stderr, "Failed to parse -d option value.\n");
                return EXIT_FAILURE;
            }
            break;
        }
        case 'c':
            fprintf(stderr, "no CPU device: this build runs its kernels on an MI355X (gfx950) GPU\n");
            return EXIT_FAILURE;
        case 'v': verbose = true; break;
        case 'M':
            tr.model = esth_model_lookup(optarg);
            if (tr.model != ESTH_MODEL_ROBOTARM5 && tr.model != ESTH_MODEL_ROBOTARM5_CAMERA &&
                tr.model != ESTH_MODEL_UNICYCLE10) {
                fprintf(stderr, "-M: robotarm5, robotarm5_camera or unicycle10 (trace-driven models)\n");
                return EXIT_FAILURE;
            }
            break;
        case 'R':
            if (!strcmp(optarg, "multinomial")) resampler = ESTH_RESAMPLE_MULTINOMIAL;
            else if (!strcmp(optarg, "systematic")) resampler = ESTH_RESAMPLE_SYSTEMATIC;
            else if (!strcmp(optarg, "alias")) resampler = ESTH_RESAMPLE_ALIAS;
            else { fprintf(stderr, "-R: multinomial, systematic or alias\n"); return EXIT_FAILURE; }
            break;
        case 'A': adaptive = true; break;
        case 's': seed = (uint32_t)strtoul(optarg, NULL, 10); break;
        case 'T': per_step = true; break;
        default: usage(argv[0]); return EXIT_FAILURE;
        }
    }
    if (optind >= argc) {
        fprintf(stderr, "no input file given\n");
        return EXIT_FAILURE;
    }
    if (m0 < 1 || g0 < 1 || t0 < 0 || loops < 0) {   /* the sweeps double m and N */
        usage(argv[0]);
        return EXIT_FAILURE;
    }
    if (adaptive) {
        resampler = ESTH_RESAMPLE_ALIAS;
        rinc = 0.01;
        if (!r_given) r0 = r1 = 0.0;
    }
    if (load_trace(argv[optind], &tr) != ESTH_OK) {
        fprintf(stderr, "could not read trace %s: %s\n", argv[optind], esth_last_error());
        free_trace(&tr);
        return EXIT_FAILURE;
    }
    esth_engine* e = NULL;
    if (esth_engine_create(device, 0, &e) != ESTH_OK) {
        fprintf(stderr, "device %d: %s\n", device, esth_last_error());
        free_trace(&tr);
        return EXIT_FAILURE;
    }
    if (verbose) {
        size_t used = 0, budget = 0;
        esth_engine_memory(e, &used, &budget);
        printf("Device: %d (gfx950)  HBM budget: %zu MiB  trace rows: %d\n", device, budget >> 20, tr.rows);
    }
    esth_unicycle10_params u10;
    esth_camera_params cam;
    const void* params = NULL;
    size_t psize = 0;
    if (tr.model == ESTH_MODEL_UNICYCLE10) {
        esth_unicycle10_params_default(&u10);
        params = &u10;
        psize = sizeof(u10);
    } else if (tr.model == ESTH_MODEL_ROBOTARM5_CAMERA) {
        esth_camera_params_default(&cam);
        params = &cam;
        psize = sizeof(cam);
    }
    printf(per_step ? "ex ey ew ee ax ay\n" : "N m r t e\n");
    int status = EXIT_SUCCESS;
    for (int t = t0; t <= t1 && status == EXIT_SUCCESS; ++t)
        for (int G = g0; G <= g1 && status == EXIT_SUCCESS; G *= 2)
            for (int m = m0; m <= m1 && status == EXIT_SUCCESS; m *= 2)
                for (double r = r0; r <= r1 && status == EXIT_SUCCESS; r += rinc)
                    for (int i = 0; i < loops; ++i) {
                        esth_filter_desc d;
                        esth_filter_desc_default(&d, tr.model);
                        d.num_particles = (uint32_t)m;
                        d.num_blocks = (uint32_t)G;
                        d.num_transfer = (uint32_t)t;
                        d.topology = t ? ESTH_TOPO_RING : ESTH_TOPO_NONE;
                        d.resampling_threshold = (float)r;
                        d.resampler = resampler;
                        d.adaptive = adaptive ? 1 : 0;
                        d.seed = seed;
                        double err = 0.0;
                        if (run(e, &d, params, psize, &tr, per_step, &err) != ESTH_OK) {
                            fprintf(stderr, "m=%d N=%d t=%d r=%f: %s\n", m, G, t, r, esth_last_error());
                            status = EXIT_FAILURE;
                            break;
                        }
                        if (!per_step) printf("%d %d %f %d %.16f\n", G, m, r, t, err);
                        fflush(stdout);
                    }
    esth_engine_destroy(e);
    free_trace(&tr);
    return status;
}
