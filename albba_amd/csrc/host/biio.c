/*
 * Graph loading with the reference's biio semantics.
 *
 *   .cbd  little-endian, no magic: int32 nrows, ncols, nnz; int32
 *         row_ptr[nrows+1]; int32 col_idx[nnz]; float64 val[nnz]
 *         (bfs/src/biio.h:21-96).
 *   .mtx  Matrix Market coordinate files (bfs/src/mmio_highlevel.h:11-299):
 *         1-based -> 0-based; symmetric/hermitian files mirror every
 *         off-diagonal entry; pattern entries become 1.0; complex entries
 *         keep the real part; within a row, entries keep file order, a
 *         mirrored entry landing right after the original's row slot.
 *
 * Deliberate differences, none of which changes a successful load: files are
 * opened "rb" (the reference opens .cbd files "r+", which fails on read-only
 * files), negative sizes and short reads return errors instead of reading
 * garbage, array-format .mtx files are rejected (-3), and the [INFO] lines of
 * bfs/src/biio.h:4-10 are printed only when ALBBA_VERBOSE is set.
 */
#include "internal.h"

#include <ctype.h>
#include <errno.h>

static int verbose(void)
{
    static int v = -1;
    if (v < 0) {
        const char *e = getenv("ALBBA_VERBOSE");
        v = (e && *e && *e != '0') ? 1 : 0;
    }
    return v;
}

#define INFO(...)                                  \
    do {                                           \
        if (verbose()) {                           \
            fprintf(stdout, "[INFO]  " __VA_ARGS__); \
        }                                          \
    } while (0)

int binary_read_csr(int *row, int *col, int *nnz, int **row_ptr, int **col_idx,
                    double **val, char *filename)
{
    FILE *fp = fopen(filename, "rb");
    if (!fp) {
        INFO("binary file %s don't exit\n", filename);
        return -1;
    }
    int hdr[3];
    if (fread(hdr, sizeof(int), 3, fp) != 3 || hdr[0] < 0 || hdr[1] < 0 || hdr[2] < 0) {
        INFO("read header error\n");
        fclose(fp);
        return -1;
    }
    *row = hdr[0];
    *col = hdr[1];
    *nnz = hdr[2];
    int *rp = (int *)malloc(sizeof(int) * ((size_t)hdr[0] + 1));
    int *ci = (int *)malloc(sizeof(int) * ((size_t)hdr[2] + 1));
    double *va = (double *)malloc(sizeof(double) * ((size_t)hdr[2] + 1));
    if (!rp || !ci || !va) {
        free(rp); free(ci); free(va);
        fclose(fp);
        return -1;
    }
    if (fread(rp, sizeof(int), (size_t)hdr[0] + 1, fp) != (size_t)hdr[0] + 1) {
        INFO("read row_ptr error\n");
        goto fail;
    }
    if (fread(ci, sizeof(int), (size_t)hdr[2], fp) != (size_t)hdr[2]) {
        INFO("read col_idx error\n");
        goto fail;
    }
    if (fread(va, sizeof(double), (size_t)hdr[2], fp) != (size_t)hdr[2]) {
        INFO("read val error\n");
        goto fail;
    }
    fclose(fp);
    *row_ptr = rp;
    *col_idx = ci;
    *val = va;
    return 0;
fail:
    free(rp); free(ci); free(va);
    fclose(fp);
    return -1;
}

void binary_write_csr(int row, int col, int nnz, int *row_ptr, int *col_idx, double *val,
                      char *filename)
{
    INFO("write binary file %s\n", filename);
    FILE *fp = fopen(filename, "wb");
    if (!fp) {
        INFO("open error\n");
        return;
    }
    INFO("write row is %d col is %d nnz is %d\n", row, col, nnz);
    fwrite(&row, sizeof(int), 1, fp);
    fwrite(&col, sizeof(int), 1, fp);
    fwrite(&nnz, sizeof(int), 1, fp);
    fwrite(row_ptr, sizeof(int), (size_t)row + 1, fp);
    fwrite(col_idx, sizeof(int), (size_t)nnz, fp);
    fwrite(val, sizeof(double), (size_t)nnz, fp);
    fclose(fp);
}

/* ---- Matrix Market --------------------------------------------------- */

typedef struct {
    char kind;       /* 'C' coordinate, 'A' array                 */
    char field;      /* 'R' real, 'C' complex, 'P' pattern, 'I' int */
    char sym;        /* 'G', 'S', 'H', 'K'                          */
    int m, n, nz;    /* size line                                   */
    char *body;      /* file contents after the size line (owned)   */
} mm_file;

static void lower(char *s)
{
    for (; *s; ++s) *s = (char)tolower((unsigned char)*s);
}

/* Returns 0, -1 (cannot open), -2 (banner), -3 (unsupported), -4 (size). */
static int mm_open(const char *filename, mm_file *mf)
{
    memset(mf, 0, sizeof(*mf));
    FILE *f = fopen(filename, "rb");
    if (!f) return -1;
    char line[1100];
    if (!fgets(line, sizeof line, f)) { fclose(f); return -2; }
    char banner[64], obj[64], fmt[64], field[64], sym[64];
    if (sscanf(line, "%63s %63s %63s %63s %63s", banner, obj, fmt, field, sym) != 5) {
        fclose(f);
        return -2;
    }
    lower(obj); lower(fmt); lower(field); lower(sym);
    if (strncmp(banner, "%%MatrixMarket", 14) != 0 || strcmp(obj, "matrix") != 0) {
        fclose(f);
        return -2;
    }
    if (!strcmp(fmt, "coordinate")) mf->kind = 'C';
    else if (!strcmp(fmt, "array")) mf->kind = 'A';
    else { fclose(f); return -2; }
    if (!strcmp(field, "real")) mf->field = 'R';
    else if (!strcmp(field, "complex")) mf->field = 'C';
    else if (!strcmp(field, "pattern")) mf->field = 'P';
    else if (!strcmp(field, "integer")) mf->field = 'I';
    else { fclose(f); return -2; }
    if (!strcmp(sym, "general")) mf->sym = 'G';
    else if (!strcmp(sym, "symmetric")) mf->sym = 'S';
    else if (!strcmp(sym, "hermitian")) mf->sym = 'H';
    else if (!strcmp(sym, "skew-symmetric")) mf->sym = 'K';
    else { fclose(f); return -2; }
    if (mf->kind != 'C') { fclose(f); return -3; }

    /* skip comment lines, then read "M N nz" */
    long body_start;
    for (;;) {
        body_start = ftell(f);
        if (!fgets(line, sizeof line, f)) { fclose(f); return -4; }
        if (line[0] != '%') break;
    }
    fseek(f, 0, SEEK_END);
    long end = ftell(f);
    size_t len = (size_t)(end - body_start);
    mf->body = (char *)malloc(len + 1);
    if (!mf->body) { fclose(f); return -4; }
    fseek(f, body_start, SEEK_SET);
    if (fread(mf->body, 1, len, f) != len) { fclose(f); free(mf->body); mf->body = NULL; return -4; }
    mf->body[len] = '\0';
    fclose(f);

    char *p = mf->body, *q;
    long v[3];
    for (int k = 0; k < 3; ++k) {
        errno = 0;
        v[k] = strtol(p, &q, 10);
        if (q == p || errno) { free(mf->body); mf->body = NULL; return -4; }
        p = q;
    }
    if (v[0] < 0 || v[1] < 0 || v[2] < 0 || v[0] > INT32_MAX || v[1] > INT32_MAX ||
        v[2] > INT32_MAX) {
        free(mf->body); mf->body = NULL;
        return -4;
    }
    mf->m = (int)v[0];
    mf->n = (int)v[1];
    mf->nz = (int)v[2];
    /* keep the entries only */
    memmove(mf->body, p, strlen(p) + 1);
    return 0;
}

/* Parse the nz entries: rows/cols 0-based, values as the reference reads them. */
static int mm_entries(mm_file *mf, int *ri, int *ci, double *va)
{
    char *p = mf->body, *q;
    for (int i = 0; i < mf->nz; ++i) {
        long r = strtol(p, &q, 10);
        if (q == p) return -4;
        p = q;
        long c = strtol(p, &q, 10);
        if (q == p) return -4;
        p = q;
        double v = 1.0;
        if (mf->field == 'R' || mf->field == 'C') {
            v = strtod(p, &q);
            if (q == p) return -4;
            p = q;
            if (mf->field == 'C') {
                (void)strtod(p, &q);
                if (q == p) return -4;
                p = q;
            }
        } else if (mf->field == 'I') {
            long iv = strtol(p, &q, 10);
            if (q == p) return -4;
            p = q;
            v = (double)(int)iv;
        }
        if (r < 1 || c < 1 || r > mf->m || c > mf->n) return -4;
        ri[i] = (int)(r - 1);
        ci[i] = (int)(c - 1);
        va[i] = v;
    }
    return 0;
}

static int mm_load(const char *filename, mm_file *mf, int **ri, int **ci, double **va)
{
    int rc = mm_open(filename, mf);
    if (rc) return rc;
    *ri = (int *)malloc(sizeof(int) * ((size_t)mf->nz + 1));
    *ci = (int *)malloc(sizeof(int) * ((size_t)mf->nz + 1));
    *va = (double *)malloc(sizeof(double) * ((size_t)mf->nz + 1));
    if (!*ri || !*ci || !*va) rc = -4;
    else rc = mm_entries(mf, *ri, *ci, *va);
    free(mf->body);
    mf->body = NULL;
    if (rc) {
        free(*ri); free(*ci); free(*va);
        *ri = *ci = NULL;
        *va = NULL;
    }
    return rc;
}

static int mirrored(const mm_file *mf) { return mf->sym == 'S' || mf->sym == 'H'; }

int mmio_info(int *m, int *n, int *nnz, int *isSymmetric, char *filename)
{
    mm_file mf;
    int *ri, *ci;
    double *va;
    int rc = mm_load(filename, &mf, &ri, &ci, &va);
    if (rc) {
        if (rc == -2) { printf("Could not process Matrix Market banner.\n"); fflush(stdout); }
        return rc;
    }
    int64_t total = mf.nz;
    if (mirrored(&mf))
        for (int i = 0; i < mf.nz; ++i)
            if (ri[i] != ci[i]) ++total;
    free(ri); free(ci); free(va);
    if (total > INT32_MAX) return -4;
    *m = mf.m;
    *n = mf.n;
    *nnz = (int)total;
    *isSymmetric = mirrored(&mf);
    return 0;
}

int mmio_data(int *csrRowPtr, int *csrColIdx, double *csrVal, char *filename)
{
    mm_file mf;
    int *ri, *ci;
    double *va;
    int rc = mm_load(filename, &mf, &ri, &ci, &va);
    if (rc) {
        if (rc == -2) { printf("Could not process Matrix Market banner.\n"); fflush(stdout); }
        return rc;
    }
    const int sym = mirrored(&mf);
    int *cnt = (int *)calloc((size_t)mf.m + 1, sizeof(int));
    if (!cnt) { free(ri); free(ci); free(va); return -4; }
    for (int i = 0; i < mf.nz; ++i) {
        cnt[ri[i]]++;
        if (sym && ri[i] != ci[i]) cnt[ci[i]]++;
    }
    csrRowPtr[0] = 0;
    for (int r = 0; r < mf.m; ++r) csrRowPtr[r + 1] = csrRowPtr[r] + cnt[r];
    memset(cnt, 0, sizeof(int) * ((size_t)mf.m + 1));
    for (int i = 0; i < mf.nz; ++i) {
        int r = ri[i], c = ci[i];
        int at = csrRowPtr[r] + cnt[r]++;
        csrColIdx[at] = c;
        csrVal[at] = va[i];
        if (sym && r != c) {
            at = csrRowPtr[c] + cnt[c]++;
            csrColIdx[at] = r;
            csrVal[at] = va[i];
        }
    }
    free(cnt); free(ri); free(ci); free(va);
    return 0;
}

static int ends_with3(const char *s, const char *ext)
{
    size_t L = strlen(s);
    return L >= 3 && s[L - 3] == ext[0] && s[L - 2] == ext[1] && s[L - 1] == ext[2];
}

static int load_mtx_alloc(int *row, int *col, int *nnz, int **row_ptr, int **col_idx,
                          double **val, int *isSym, char *filename)
{
    int rc = mmio_info(row, col, nnz, isSym, filename);
    if (rc) return rc;
    *row_ptr = (int *)malloc(sizeof(int) * ((size_t)*row + 1));
    *col_idx = (int *)malloc(sizeof(int) * ((size_t)*nnz + 1));
    *val = (double *)malloc(sizeof(double) * ((size_t)*nnz + 1));
    if (!*row_ptr || !*col_idx || !*val) return -4;
    return mmio_data(*row_ptr, *col_idx, *val, filename);
}

void read_Dmatrix_convert(int *row, int *col, int *nnz, int **row_ptr, int **col_idx,
                          double **val, int *isSymmetric, char *filename)
{
    size_t L = strlen(filename);
    if (L < 3 || L > 500) return;
    char binary_name[512];
    memcpy(binary_name, filename, L + 1);
    binary_name[L - 1] = 'd';
    binary_name[L - 2] = 'b';
    binary_name[L - 3] = 'c';
    if (binary_read_csr(row, col, nnz, row_ptr, col_idx, val, binary_name) == -1) {
        INFO("read file name is %s\n", filename);
        if (load_mtx_alloc(row, col, nnz, row_ptr, col_idx, val, isSymmetric, filename) == 0) {
            INFO("begin write binary file\n");
            binary_write_csr(*row, *col, *nnz, *row_ptr, *col_idx, *val, binary_name);
            INFO("end write binary file\n");
        }
    } else {
        INFO("read binary file %s\n", binary_name);
    }
}

void read_Dmatrix(int *row, int *col, int *nnz, int **row_ptr, int **col_idx, double **val,
                  int *isSymmetric, char *filename)
{
    if (ends_with3(filename, "cbd")) {
        INFO("read binary file %s\n", filename);
        binary_read_csr(row, col, nnz, row_ptr, col_idx, val, filename);
    } else if (ends_with3(filename, "mtx")) {
        INFO("read regular file %s\n", filename);
        load_mtx_alloc(row, col, nnz, row_ptr, col_idx, val, isSymmetric, filename);
    } else {
        printf("[INFO]  File type unsupport\n");
        fflush(stdout);
    }
}
