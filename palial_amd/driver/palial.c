/*
 * palial -- command-line driver with the reference's interface
 * (palial_src/main.c:57-170):
 *
 *     palial -a {cholesky,qr,lu} -m N -b B
 *
 * generates the reference's test matrix (include/populate.h:35-66, seeded
 * here instead of srand(time(NULL))) in layout L on the host, factors it in
 * place on the GPU through libpalial_hip.so and stays silent on stdout, as
 * the reference does.  Same messages and exit status for missing arguments
 * and for N % B != 0.
 *
 * Additive options (the reference has none of these):
 *   -s SEED   generator seed (default 0)
 *   -g GEN    palial (default) | sym (S + N I) | dense (G + N I)
 *   -t        print "algorithm, N, B, factor_ms, TFLOP/s" after the run
 *   -v        check the result: Cholesky/LU solve one right-hand side with
 *             the factorization tier and print the backward error
 *   -T        trace the factorization per task: a loaded tracer that defines
 *             the reference's hooks (include/common.h:9-11) receives each
 *             task's name and device time
 * Library errors go to stderr with exit status 1.
 */
#include "palial_hip.h"

#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int fail(int rc, const char *what)
{
    fprintf(stderr, "palial: %s failed: %s (%s)\n", what, palial_hip_strerror(rc), palial_hip_last_error());
    return 1;
}

int main(int argc, char *argv[])
{
    char algorithm[16] = "";
    int N = 0, b = 0, timing = 0, verify = 0, trace = 0;
    unsigned long long seed = 0;
    int gen = PALIAL_GEN_PALIAL;
    static struct option longopts[] = {
        {"Algorithm", required_argument, NULL, 'a'},
        {"Matrix size", required_argument, NULL, 'm'},
        {"Tile size", required_argument, NULL, 'b'},
        {NULL, 0, NULL, 0},
    };
    if (argc < 2) {
        printf("Missing arguments.\n");
        printf("Aborting...\n");
        return 0;
    }
    int c;
    while ((c = getopt_long(argc, argv, "a:m:b:s:g:tvTh", longopts, NULL)) != -1) {
        switch (c) {
        case 'a': snprintf(algorithm, sizeof algorithm, "%s", optarg); break;
        case 'm': N = atoi(optarg); break;
        case 'b': b = atoi(optarg); break;
        case 's': seed = strtoull(optarg, NULL, 10); break;
        case 'g':
            if (!strcmp(optarg, "sym")) gen = PALIAL_GEN_SYM;
            else if (!strcmp(optarg, "dense")) gen = PALIAL_GEN_DENSE;
            else gen = PALIAL_GEN_PALIAL;
            break;
        case 't': timing = 1; break;
        case 'v': verify = 1; break;
        case 'T': trace = 1; break;
        case 'h': printf("HELP\n"); return 0;
        default:
            printf("Invalid arguments.\n");
            printf("Aborting...\n");
            return 0;
        }
    }
    if (b <= 0 || N <= 0 || N % b != 0) {
        printf("Tile size does not divide the matrix size. Aborting...\n");
        return 0;
    }
    int algo = !strcmp(algorithm, "cholesky") ? 0 : !strcmp(algorithm, "lu") ? 1 : !strcmp(algorithm, "qr") ? 2 : -1;
    /* an unknown algorithm name runs nothing and prints nothing, as the
     * reference's if / else-if chain (main.c:123-158) */
    if (algo < 0) return 0;

    int rc = palial_hip_init(0);
    if (rc) return fail(rc, "init");
    if (trace && (rc = palial_hip_trace(1))) return fail(rc, "trace");
    const size_t n2 = (size_t)N * N;
    double *a = NULL, *s = NULL;
    if (posix_memalign((void **)&a, 4096, n2 * sizeof(double))) {
        printf("Problem allocating contiguous memory.\n");
        return 1;
    }
    MATRIX_desc A = {b, (long)n2, (long)b * b, N, a};
    if ((rc = palial_hip_generate(A, gen, seed, (double)N, NULL))) return fail(rc, "generate");
    double *a0 = NULL;
    if (verify && algo < 2) {
        a0 = malloc(n2 * sizeof(double));
        if (!a0) return fail(PALIAL_ERR_NOMEM, "allocation");
        memcpy(a0, a, n2 * sizeof(double));
    }

    palial_report rep;
    memset(&rep, 0, sizeof rep);
    if (algo == 0) {
        rc = palial_hip_cholesky(A, NULL, &rep);
    } else if (algo == 1) {
        rc = palial_hip_lu(A, NULL, NULL, &rep);
    } else {
        /* workspace S with the same descriptor shape (main.c:137-148) */
        s = calloc(n2, sizeof(double));
        if (!s) return fail(PALIAL_ERR_NOMEM, "allocation");
        MATRIX_desc S = {b, (long)n2, (long)b * b, N, s};
        rc = palial_hip_qr(A, S, NULL, &rep);
    }
    if (rc < 0) return fail(rc, algorithm);
    if (rc > 0) fprintf(stderr, "palial: %s: non-positive pivot at column %d\n", algorithm, rc);
    if (timing) printf("%s, %d, %d, %.3f, %.3f\n", algorithm, N, b, rep.factor_ms, rep.tflops);

    if (verify && a0) {
        double *x = malloc((size_t)N * sizeof(double));
        if (!x) return fail(PALIAL_ERR_NOMEM, "allocation");
        for (int i = 0; i < N; i++) x[i] = 1.0;
        MATRIX_desc A0 = {b, (long)n2, (long)b * b, N, a0};
        palial_report r2;
        memset(&r2, 0, sizeof r2);
        rc = palial_hip_solve(algo, A0, x, 1, N, NULL, &r2);
        if (rc) return fail(rc, "solve");
        printf("backward error %.3e after %d refinement steps\n", r2.backward_error, r2.iterations);
        free(x);
        free(a0);
    }
    free(s);
    free(a);
    palial_hip_finalize();
    return 0;
}
