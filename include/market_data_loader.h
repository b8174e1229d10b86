/*
 * Return series for the filter: the reference's market-data formats
 * (include/market_data_loader.h:37-237), with the same function names,
 * arguments, allocation (malloc, freed by the caller) and results.
 *
 *   <name>.bin           raw little-endian float32 returns, no header
 *   <name>_returns.csv   one return per line
 *   <name>.csv           header line, then date,price,return,log_return
 *
 * Header-only C, host side.  The Python twins are svpf_amd.data.load_returns_binary,
 * load_returns_csv, load_market_data_csv and return_stats.
 */
#ifndef SVPF_MARKET_DATA_LOADER_H
#define SVPF_MARKET_DATA_LOADER_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/market_data_loader.h:37-67: *n = file size / 4 floats; 0 on
 * success, -1 when the file cannot be opened, allocated or fully read. */
static inline int load_returns_binary(const char* filepath, float** returns, int* n) {
    FILE* f = fopen(filepath, "rb");
    if (!f) {
        fprintf(stderr, "ERROR: Cannot open %s\n", filepath);
        return -1;
    }
    long bytes = -1;
    if (fseek(f, 0, SEEK_END) == 0) bytes = ftell(f);
    if (bytes < 0 || fseek(f, 0, SEEK_SET) != 0) {
        fclose(f);
        return -1;
    }
    const int count = (int)(bytes / (long)sizeof(float));
    float* buf = (float*)malloc(bytes > 0 ? (size_t)bytes : 1);
    if (!buf) {
        fclose(f);
        return -1;
    }
    const size_t got = fread(buf, sizeof(float), (size_t)count, f);
    fclose(f);
    if (got != (size_t)count) {
        free(buf);
        *returns = NULL;
        return -1;
    }
    *returns = buf;
    *n = count;
    return 0;
}

/* include/market_data_loader.h:77-112: every line (up to 255 characters) is
 * one value, parsed with atof (a blank line reads as 0). */
static inline int load_returns_csv(const char* filepath, float** returns, int* n) {
    FILE* f = fopen(filepath, "r");
    if (!f) {
        fprintf(stderr, "ERROR: Cannot open %s\n", filepath);
        return -1;
    }
    char line[256];
    int count = 0;
    while (fgets(line, sizeof line, f)) ++count;
    float* buf = (float*)malloc(count > 0 ? (size_t)count * sizeof(float) : 1);
    if (!buf) {
        fclose(f);
        return -1;
    }
    rewind(f);
    for (int i = 0; i < count && fgets(line, sizeof line, f); ++i) buf[i] = (float)atof(line);
    fclose(f);
    *returns = buf;
    *n = count;
    return 0;
}

/* include/market_data_loader.h:123-180: skips the header line; returns the
 * fourth column (log_return) and, when `prices` is not NULL, the second. */
static inline int load_market_data_csv(const char* filepath, float** returns, float** prices, int* n) {
    FILE* f = fopen(filepath, "r");
    if (!f) {
        fprintf(stderr, "ERROR: Cannot open %s\n", filepath);
        return -1;
    }
    char line[512];
    if (!fgets(line, sizeof line, f)) {
        fclose(f);
        return -1;
    }
    int count = 0;
    while (fgets(line, sizeof line, f)) ++count;
    float* r = (float*)malloc(count > 0 ? (size_t)count * sizeof(float) : 1);
    float* p = prices ? (float*)malloc(count > 0 ? (size_t)count * sizeof(float) : 1) : NULL;
    if (!r || (prices && !p)) {
        free(r);
        free(p);
        fclose(f);
        return -1;
    }
    rewind(f);
    if (!fgets(line, sizeof line, f)) count = 0;
    for (int i = 0; i < count && fgets(line, sizeof line, f); ++i) {
        /* date,price,return,log_return */
        const char* col[4] = {line, NULL, NULL, NULL};
        int k = 1;
        for (char* c = line; *c && k < 4; ++c)
            if (*c == ',') col[k++] = c + 1;
        r[i] = col[3] ? (float)atof(col[3]) : 0.0f;
        if (p) p[i] = col[1] ? (float)atof(col[1]) : 0.0f;
    }
    fclose(f);
    *returns = r;
    if (prices) *prices = p;
    *n = count;
    return 0;
}

/* include/market_data_loader.h:185-222: mean and population std summed in
 * double, min / max, and the counts of |r - mean| / std above 3 and 4. */
static inline void compute_return_stats(const float* returns, int n, float* mean, float* std, float* min_ret,
                                        float* max_ret, int* n_3sigma, int* n_4sigma) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += returns[i];
    const float m = (float)(s / n);
    double ss = 0.0;
    for (int i = 0; i < n; ++i) {
        const double d = returns[i] - m;
        ss += d * d;
    }
    const float sd = (float)sqrt(ss / n);
    float lo = returns[0], hi = returns[0];
    int c3 = 0, c4 = 0;
    for (int i = 0; i < n; ++i) {
        lo = returns[i] < lo ? returns[i] : lo;
        hi = returns[i] > hi ? returns[i] : hi;
        const float z = fabsf((returns[i] - m) / sd);
        c3 += z > 3.0f;
        c4 += z > 4.0f;
    }
    *mean = m;
    *std = sd;
    *min_ret = lo;
    *max_ret = hi;
    *n_3sigma = c3;
    *n_4sigma = c4;
}

/* include/market_data_loader.h:227-237: annualised (252 days) summary */
static inline void print_return_stats(const char* name, const float* returns, int n) {
    float m, sd, lo, hi;
    int c3, c4;
    compute_return_stats(returns, n, &m, &sd, &lo, &hi, &c3, &c4);
    printf("  %s: n=%d, mean=%.4f, std=%.4f, range=[%.2f%%, %.2f%%]\n", name, n, m * 252, sd * sqrtf(252.0f),
           lo * 100, hi * 100);
    printf("    Tail events: %d (>3 std), %d (>4 std)\n", c3, c4);
}

#ifdef __cplusplus
}
#endif

#endif /* SVPF_MARKET_DATA_LOADER_H */
