/* C caller of libattgnn.so: the header compiles as C99, the library loads
 * with dlopen, and the host preprocessing entry points run without a GPU.
 * Usage: abi_smoke <path to libattgnn.so>; prints the window-block arrays of
 * a 12-row CSR (window 8, wide 4, part size 2, fp32 degree products). */
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>

#include "attgnn.h"

typedef int (*build_fn)(const int32_t *, int64_t, const int32_t *, int64_t, int, const void *, int, int, int, int,
                        int, int, int, attgnn_blocks **);
typedef int (*sizes_fn)(const attgnn_blocks *, int64_t *, int64_t *, int64_t *, int64_t *, int64_t *, int *);
typedef int (*export_fn)(const attgnn_blocks *, int32_t *, int32_t *, void *, int32_t *, int32_t *, int32_t *);
typedef void (*free_fn)(attgnn_blocks *);
typedef const char *(*status_fn)(int);
typedef int (*spmm_fn)(const int32_t *, const int32_t *, const void *, int, const int32_t *, const int32_t *, int64_t,
                       const void *, int, int64_t, void *, int, int64_t, int64_t, int, int, int, const float *,
                       const float *, float *, int64_t, int, attgnn_stream_t);

static void *sym(void *h, const char *name)
{
    void *p = dlsym(h, name);
    if (!p) {
        fprintf(stderr, "missing symbol %s\n", name);
        exit(2);
    }
    return p;
}

static void print_i32(const char *name, const int32_t *a, int64_t n)
{
    printf("%s", name);
    for (int64_t i = 0; i < n; ++i) printf(" %d", a[i]);
    printf("\n");
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    void *h = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        fprintf(stderr, "dlopen: %s\n", dlerror());
        return 2;
    }
    build_fn build = (build_fn)sym(h, "attgnn_blocks_build");
    sizes_fn sizes = (sizes_fn)sym(h, "attgnn_blocks_sizes");
    export_fn exportb = (export_fn)sym(h, "attgnn_blocks_export");
    free_fn freeb = (free_fn)sym(h, "attgnn_blocks_free");
    status_fn status = (status_fn)sym(h, "attgnn_status_string");
    spmm_fn spmm = (spmm_fn)sym(h, "attgnn_spmm_block");

    /* 12 rows; row 5 empty; unsorted row 2 */
    const int32_t row[13] = {0, 3, 4, 7, 8, 10, 10, 12, 13, 15, 16, 18, 19};
    const int32_t col[19] = {0, 4, 9, 1, 7, 2, 5, 3, 0, 11, 6, 8, 7, 2, 10, 9, 1, 11, 0};
    float d[12];
    for (int i = 0; i < 12; ++i) d[i] = 1.0f / (float)(i + 1);

    attgnn_blocks *b = NULL;
    int rc = build(row, 13, col, 19, ATTGNN_SRC_DEG_F32, d, ATTGNN_F32, 8, 4, 2, 1, 0, 1, &b);
    if (rc != ATTGNN_OK) {
        fprintf(stderr, "build: %s\n", status(rc));
        return 1;
    }
    int64_t n_rp, n_col, n_val, n_parts, n_unique;
    int has_t;
    sizes(b, &n_rp, &n_col, &n_val, &n_parts, &n_unique, &has_t);
    int32_t *rp = malloc(sizeof(int32_t) * n_rp), *c = malloc(sizeof(int32_t) * n_col);
    int32_t *w = malloc(sizeof(int32_t) * n_parts), *at = malloc(sizeof(int32_t) * n_parts);
    float *val = malloc(sizeof(float) * n_val);
    exportb(b, rp, c, val, NULL, w, at);
    print_i32("rp", rp, n_rp);
    print_i32("col", c, n_col);
    print_i32("window", w, n_parts);
    print_i32("atomic", at, n_parts);
    printf("val");
    for (int64_t i = 0; i < n_val; ++i) printf(" %.9g", val[i]);
    printf("\nunique %lld\n", (long long)n_unique);
    freeb(b);

    /* errors are status codes, not crashes: bad window / malformed CSR / bad spmm args */
    attgnn_blocks *bad = NULL;
    printf("err_window %d\n", build(row, 13, col, 19, ATTGNN_SRC_DEG_F32, d, ATTGNN_F32, 0, 4, 2, 1, 0, 1, &bad));
    const int32_t badrow[3] = {0, 2, 1};
    printf("err_format %d\n", build(badrow, 3, col, 1, ATTGNN_SRC_DEG_F32, d, ATTGNN_F32, 8, 4, 2, 1, 0, 1, &bad));
    printf("err_spmm %d\n", spmm(NULL, NULL, NULL, ATTGNN_F16, NULL, NULL, 0, NULL, ATTGNN_F16, 0, NULL, ATTGNN_F32, 0,
                                 16, 12, 8, ATTGNN_MODE_NATIVE, NULL, NULL, NULL, 0, 0, NULL));
    free(rp); free(c); free(w); free(at); free(val);
    /* no dlclose: the HIP runtime the library links is not meant to be unloaded */
    return 0;
}
