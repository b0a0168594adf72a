/*
 * compute_e_main.c — command-line driver: decimal digits of e on the GPU.
 *
 * Usage and output follow the reference's compute_e
 * (ntt/app/compute_e/main.c:155-165 arguments, :167-173 header line,
 * :340-345 the "e = 2.ddd..." line with D significant digits):
 *
 *   compute_e -d <digits> [-f|--force] [--plain] [-t tuning.json]
 *
 * -t loads a tuning file (ntt_engine_autotune_save's format) after init; by
 * default the file tuning_mi355x.json beside the executable is loaded when it
 * exists (the NTT_AMD_TUNING environment variable, read at init, takes
 * precedence over both).
 * --plain evaluates the product tree with the reference's depth-first
 * recursion (one multiply call per product) instead of the level-batched
 * device tree.  -f is accepted for compatibility; there is no GPU-time gate.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/binary_split.h"
#include "../../include/compute_e.h"
#include "../../include/ntt_engine.h"

int main(int argc, char **argv)
{
    long n_digits = 1000;
    int batched = 1;
    const char *tuning = NULL;
    for (int i = 1; i < argc; i++) {
        if (strcmp(argv[i], "-d") == 0 && i + 1 < argc) {
            n_digits = atol(argv[++i]);
            if (n_digits < 1) {
                fprintf(stderr, "digits must be >= 1\n");
                return 1;
            }
        } else if (strcmp(argv[i], "-f") == 0 || strcmp(argv[i], "--force") == 0) {
            /* accepted for compatibility */
        } else if (strcmp(argv[i], "--plain") == 0) {
            batched = 0;
        } else if (strcmp(argv[i], "-t") == 0 && i + 1 < argc) {
            tuning = argv[++i];
        } else {
            fprintf(stderr, "Usage: %s [-d digits] [-f/--force] [--plain] [-t tuning.json]\n", argv[0]);
            return 1;
        }
    }
    long n_guard = 20;
    long n_terms = split_terms_needed(n_digits + n_guard);
    printf("compute_e: %ld digits, %ld terms, %ld guard\n", n_digits, n_terms, n_guard);
    fflush(stdout);
    int rc = ntt_engine_init(0);
    if (rc) {
        fprintf(stderr, "compute_e: %s\n", ntt_engine_last_error());
        return 1;
    }
    /* the tuning file: -t, else tuning_mi355x.json beside the executable */
    char beside[4096];
    if (!tuning && !getenv("NTT_AMD_TUNING")) {
        const char *slash = strrchr(argv[0], '/');
        const int dir = slash ? (int)(slash - argv[0]) + 1 : 0;
        snprintf(beside, sizeof beside, "%.*stuning_mi355x.json", dir, argv[0]);
        FILE *f = fopen(beside, "r");
        if (f) {
            fclose(f);
            tuning = beside;
        }
    }
    if (tuning && (rc = ntt_engine_autotune_load(tuning)) != 0) {
        fprintf(stderr, "compute_e: %s\n", ntt_engine_last_error());
        return 1;
    }
    char *out = (char *)malloc((size_t)n_digits + 2);
    if (!out) {
        fprintf(stderr, "compute_e: out of memory\n");
        return 1;
    }
    ntt_e_stats st;
    rc = ntt_compute_e(n_digits, batched, out, (size_t)n_digits + 2, &st);
    if (rc) {
        fprintf(stderr, "compute_e: %s (status %d)\n", ntt_engine_last_error(), rc);
        return 1;
    }
    printf("  split done: %.2f s (%s tree)\n", st.t_tree, batched ? "level-batched" : "depth-first");
    printf("  reciprocal: %.2f s\n", st.t_recip);
    printf("  division done: %.2f s\n", st.t_div);
    printf("  base conversion done: %.2f s\n", st.t_radix);
    printf("\ne = ");
    size_t len = (size_t)n_digits + 1, wr = fwrite(out, 1, len, stdout);
    wr += fwrite("\n", 1, 1, stdout);
    if (wr != len + 1) {
        fprintf(stderr, "compute_e: short write\n");
        return 1;
    }
    printf("total: %.2f s\n", st.t_total);
    printf("gpu_dispatches: %lu\n", st.products);
    free(out);
    ntt_engine_teardown();
    return 0;
}
