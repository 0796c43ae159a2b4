/* LinA's sparse-matrix files (linear_algebra.c:832-916, 983-1002, 1023-1056):
 *
 * - binary (write_sparse / read_sparse): int n, int Len[n], then every row's
 *   column indices (int), then every row's values (double), native byte
 *   order, no header or padding;
 * - text (write_sparse_textfile / read_sparse_textfile): one line per row,
 *   entries "(j|v)" in row order, two spaces between entries, values printed
 *   with "%f";
 * - MatrixMarket (print_to_matrix_market_format): "<Name>.mtx", coordinate
 *   real general, the nonzero entries only, 1-based, "%g".
 *
 * Files written here are byte-identical to the reference's for the same row
 * lists, and either library reads the other's files to the same matrix. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dropin.h"

static FILE* open_in(const char* path, const char* name, const char* mode) {
    char full[512];
    snprintf(full, sizeof(full), "%s/%s", path, name);
    return fopen(full, mode);
}

/* write_sparse (linear_algebra.c:832-844): 1 on success, 0 on a failed
 * write (or a file that cannot be opened) */
int write_sparse(sparse_matrix* A, char* path, char* name) {
    int i, n = A->size, ok = 1;
    FILE* f = open_in(path, name, "w");
    if (!f) return 0;
    if (fwrite(&n, sizeof(int), 1, f) != 1) ok = 0;
    if (ok && fwrite(A->Len, sizeof(int), n, f) != (size_t)n) ok = 0;
    for (i = 0; ok && i < n; i++)
        if (fwrite(A->Indices[i], sizeof(int), A->Len[i], f) != (size_t)A->Len[i]) ok = 0;
    for (i = 0; ok && i < n; i++)
        if (fwrite(A->Values[i], sizeof(double), A->Len[i], f) != (size_t)A->Len[i]) ok = 0;
    fclose(f);
    return ok;
}

/* read_sparse (linear_algebra.c:846-868): NULL on a missing or short file */
sparse_matrix* read_sparse(char* path, char* name) {
    int i, n;
    sparse_matrix* R;
    FILE* f = open_in(path, name, "r");
    if (!f) return NULL;
    if (fread(&n, sizeof(int), 1, f) != 1 || n < 0) {
        fclose(f);
        return NULL;
    }
    R = (sparse_matrix*)malloc(sizeof(sparse_matrix));
    R->size = n;
    R->Len = (int*)malloc((n > 0 ? n : 1) * sizeof(int));
    R->Indices = (int**)calloc(n > 0 ? n : 1, sizeof(int*));
    R->Values = (double**)calloc(n > 0 ? n : 1, sizeof(double*));
    if (fread(R->Len, sizeof(int), n, f) != (size_t)n) goto fail;
    for (i = 0; i < n; i++) {
        if (R->Len[i] < 0) goto fail;
        R->Indices[i] = (int*)malloc((R->Len[i] > 0 ? R->Len[i] : 1) * sizeof(int));
        if (fread(R->Indices[i], sizeof(int), R->Len[i], f) != (size_t)R->Len[i]) goto fail;
    }
    for (i = 0; i < n; i++) {
        R->Values[i] = (double*)malloc((R->Len[i] > 0 ? R->Len[i] : 1) * sizeof(double));
        if (fread(R->Values[i], sizeof(double), R->Len[i], f) != (size_t)R->Len[i]) goto fail;
    }
    fclose(f);
    return R;
fail:
    fclose(f);
    for (i = 0; i < n; i++) {
        free(R->Indices[i]);
        free(R->Values[i]);
    }
    free(R->Len);
    free(R->Indices);
    free(R->Values);
    free(R);
    return NULL;
}

/* one line without its '\n' into *buf (grown as needed); 1 when the line
 * ended with '\n', 0 at end of file (readLine, File_stuff.c:171-187: a last
 * line without '\n' is not a row) */
static int read_line(FILE* f, char** buf, size_t* cap) {
    size_t k = 0;
    int c;
    for (;;) {
        c = fgetc(f);
        if (c == '\n' || c == EOF) break;
        if (k + 1 >= *cap) {
            *cap = *cap ? 2 * *cap : 256;
            *buf = (char*)realloc(*buf, *cap);
        }
        (*buf)[k++] = (char)c;
    }
    if (*cap == 0) {
        *cap = 1;
        *buf = (char*)malloc(1);
    }
    (*buf)[k] = '\0';
    return c != EOF;
}

/* read_sparse_textfile (linear_algebra.c:870-916): row i = line i; the line
 * splits at spaces into tokens, a token that splits at '|' into exactly two
 * parts is an entry: the first part without its first character is the
 * column (atoi), the second without its last character the value (atof),
 * added with insert_sparse to sparse_zero's row.  Lines are read whole (the
 * reference copies at most 2048 characters of a line into its splitter). */
sparse_matrix* read_sparse_textfile(char* path, char* name) {
    FILE* f = open_in(path, name, "r");
    char *line = NULL, *tok, *save;
    size_t cap = 0;
    int rows = 0, i;
    sparse_matrix* R;
    if (!f) return NULL;
    while (read_line(f, &line, &cap)) rows++;
    rewind(f);
    R = sparse_zero(rows);
    for (i = 0; i < rows; i++) {
        read_line(f, &line, &cap);
        for (tok = strtok_r(line, " ", &save); tok; tok = strtok_r(NULL, " ", &save)) {
            char *p0, *p1, *p2, *s2;
            char* t = strdup(tok);
            p0 = strtok_r(t, "|", &s2);
            p1 = p0 ? strtok_r(NULL, "|", &s2) : NULL;
            p2 = p1 ? strtok_r(NULL, "|", &s2) : NULL;
            if (p0 && p1 && !p2) {
                size_t l1 = strlen(p1);
                p1[l1 - 1] = '\0';
                insert_sparse(R, atof(p1), i, atoi(p0 + 1));
            }
            free(t);
        }
    }
    free(line);
    fclose(f);
    return R;
}

/* write_sparse_textfile (linear_algebra.c:983-1002); a row without entries
 * is an empty line (the reference's container never holds one) */
void write_sparse_textfile(char* path, char* name, sparse_matrix* A) {
    int i, j;
    FILE* f;
    if (A == NULL) {
        printf("Matrix = NULL");
        return;
    }
    f = open_in(path, name, "w");
    if (!f) return;
    for (i = 0; i < A->size; i++) {
        const int n = A->Len[i] - 1;
        for (j = 0; j < n; j++) fprintf(f, "(%d|%f)  ", A->Indices[i][j], A->Values[i][j]);
        if (n >= 0) fprintf(f, "(%d|%f)\n", A->Indices[i][n], A->Values[i][n]);
        else fputc('\n', f);
    }
    fclose(f);
}

/* print_to_matrix_market_format (linear_algebra.c:1023-1056): the header
 * counts rows and columns up to the last nonzero entry */
void print_to_matrix_market_format(sparse_matrix* A, char* Name) {
    char full[512];
    int i, j, rows = 0, cols = 0, count = 0;
    FILE* f;
    if (A == NULL) return;
    snprintf(full, sizeof(full), "%s.mtx", Name);
    f = fopen(full, "w");
    if (!f) return;
    for (i = 0; i < A->size; i++) {
        int any = 0;
        for (j = 0; j < A->Len[i]; j++)
            if (A->Values[i][j] != 0) {
                count++;
                any = 1;
                if (A->Indices[i][j] + 1 > cols) cols = A->Indices[i][j] + 1;
            }
        if (any && i + 1 > rows) rows = i + 1;
    }
    fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    fprintf(f, "%d %d %d\n", rows, cols, count);
    for (i = 0; i < A->size; i++)
        for (j = 0; j < A->Len[i]; j++)
            if (A->Values[i][j] != 0) fprintf(f, "%d %d %g\n", i + 1, A->Indices[i][j] + 1, A->Values[i][j]);
    fclose(f);
}
