/*
 * fastfib - the reference's command line (src/main.c:28-106) over the GPU
 * library's C-ABI (include/fastfib_amd.h):
 *
 *     fastfib [n] [--print] [--help]
 *
 * Same argument rules (options in any order, one n, anything else an error),
 * the same strict n (decimal digits only, must fit in 64 bits), the same
 * stdin prompt when n is missing (a UTF-8 byte order mark and surrounding
 * blanks are skipped), the same output lines and exit codes (0, or 1 with a
 * message for bad input). F(n) and its decimal digits are computed on the
 * GPU by ff_fibonacci and ff_big_to_string. "Time:" covers F(n) only, like
 * the reference's; the device is opened (ff_init) before the clock starts,
 * as process start-up, since the reference has no device to open.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "fastfib_amd.h"

static void usage(FILE *f) {
    fputs("usage: fastfib [n] [--print]\n"
          "  n        which Fibonacci number (asked for if missing)\n"
          "  --print  also print its decimal digits\n",
          f);
}

/* 1 if s is a non-empty string of decimal digits whose value fits in 64
   bits (then *n gets it), else 0: no sign, no blanks, no exponent, no hex. */
static int read_u64(const char *s, uint64_t *n) {
    if (*s == '\0') return 0;
    uint64_t v = 0;
    for (; *s; s++) {
        if (*s < '0' || *s > '9') return 0;
        const uint64_t d = (uint64_t)(*s - '0');
        if (v > (UINT64_MAX - d) / 10) return 0; /* v * 10 + d would pass 2^64 - 1 */
        v = v * 10 + d;
    }
    *n = v;
    return 1;
}

static double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char **argv) {
    uint64_t n = 0;
    int have_n = 0, print = 0;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (!strcmp(a, "--print")) {
            print = 1;
        } else if (!strcmp(a, "--help") || !strcmp(a, "-h")) {
            usage(stdout);
            return 0;
        } else if (!have_n && read_u64(a, &n)) {
            have_n = 1;
        } else {
            fprintf(stderr, "invalid argument: %s\n", a);
            usage(stderr);
            return 1;
        }
    }
    if (!have_n) {
        char line[64];
        fputs("Enter n: ", stdout);
        fflush(stdout);
        if (!fgets(line, sizeof line, stdin)) return 1;
        char *s = line;
        if (!strncmp(s, "\xEF\xBB\xBF", 3)) s += 3; /* byte order mark of piped UTF-8 */
        s += strspn(s, " \t");
        s[strcspn(s, " \t\r\n")] = '\0';
        if (!read_u64(s, &n)) {
            fprintf(stderr, "invalid n: %s\n", s);
            return 1;
        }
    }

    if (ff_init(-1) != FF_OK) {
        fprintf(stderr, "fastfib: %s\n", ff_last_error());
        return 1;
    }
    double t0 = now_seconds();
    Big f = ff_fibonacci(n);
    const double t_fib = now_seconds() - t0;

    double t_dec = 0;
    if (print) {
        t0 = now_seconds();
        char *s = ff_big_to_string(&f);
        t_dec = now_seconds() - t0;
        printf("fibonacci(%llu) = %s\n", (unsigned long long)n, s);
        free(s);
    }
    printf("Time: %.6f seconds\n", t_fib);
    if (print) printf("Decimal conversion: %.6f seconds\n", t_dec);
    ff_big_free(&f);
    return 0;
}
