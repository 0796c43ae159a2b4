/* Crack's checkpoint file data.bin (host C of the LinA drop-in), so that a
 * run of this library and a Crack run exchange their fields: the layout of
 * init_binary_save / binary_save / binary_load (Crack/crack.c:5095-5287):
 *
 *   int n            fields per node (degrees of freedom + added fields)
 *   int m            number of records
 *   m records of     double t, int mesh_index (1-based),
 *                    n * mesh_size[mesh_index - 1] doubles (field-major)
 *
 * in the machine's byte order with no padding.  The mesh sizes come from the
 * first token of <dir>/meshes/<base>.<k>.node, k = 1, 2, ... (load_mesh_sizes,
 * :5168-5230).  Functions return SUCCESS (1) or FAILED (0) like Crack's. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dropin.h"
#include "../../include/linalg/checkpoint.h"

int lina_checkpoint_init(const char* path, int n_fields, const char* const* names) { /* :5095-5121 */
    const int m = 0;
    FILE* f = fopen(path, "wb");
    if (!f) {
        printf("could not create checkpoint file %s\n", path);
        return 0;
    }
    fwrite(&n_fields, sizeof(int), 1, f);
    fwrite(&m, sizeof(int), 1, f);
    fclose(f);
    if (names) {
        int i;
        char* info = (char*)malloc(strlen(path) + 6);
        sprintf(info, "%s.info", path);
        f = fopen(info, "w");
        free(info);
        if (!f) return 0;
        fprintf(f, "%d\n", n_fields);
        for (i = 0; i < n_fields; i++) fprintf(f, "%s\n", names[i]);
        fclose(f);
    }
    return 1;
}

/* :5124-5166: the header's record count + 1 (n rewritten), the record at the end */
int lina_checkpoint_append(const char* path, int n_fields, const double* X, int size, double t, int mesh_index) {
    int n, m;
    FILE* f = fopen(path, "r+b");
    if (!f) {
        printf("could not write to checkpoint file %s\n", path);
        return 0;
    }
    if (fread(&n, sizeof(int), 1, f) != 1 || fread(&m, sizeof(int), 1, f) != 1) {
        fclose(f);
        printf("checkpoint file %s has no header\n", path);
        return 0;
    }
    m++;
    n = n_fields;
    fseek(f, 0, SEEK_SET);
    fwrite(&n, sizeof(int), 1, f);
    fwrite(&m, sizeof(int), 1, f);
    fseek(f, 0, SEEK_END);
    fwrite(&t, sizeof(double), 1, f);
    fwrite(&mesh_index, sizeof(int), 1, f);
    if (size > 0) fwrite(X, sizeof(double), (size_t)size, f);
    fclose(f);
    return 1;
}

int lina_checkpoint_info(const char* path, int* n_fields, int* records) {
    int hdr[2];
    FILE* f = fopen(path, "rb");
    if (!f) return 0;
    if (fread(hdr, sizeof(int), 2, f) != 2) {
        fclose(f);
        return 0;
    }
    fclose(f);
    if (n_fields) *n_fields = hdr[0];
    if (records) *records = hdr[1];
    return 1;
}

/* :5168-5230 (the mesh base name given instead of read from the run's
 * parameter file) */
int lina_checkpoint_mesh_sizes(const char* dir, const char* mesh_base, int** sizes, int* count) {
    int k = 0;
    *sizes = NULL;
    *count = 0;
    for (;;) {
        char name[1024], line[1024];
        FILE* f;
        snprintf(name, sizeof name, "%s/meshes/%s.%d.node", dir, mesh_base, k + 1);
        f = fopen(name, "r");
        if (!f) break;
        k++;
        if (fgets(line, sizeof line, f)) {
            char* tok = strtok(line, " \t\n");
            if (tok) {
                *sizes = (int*)realloc(*sizes, (size_t)(*count + 1) * sizeof(int));
                (*sizes)[(*count)++] = atoi(tok);
            }
        }
        fclose(f);
    }
    return *count > 0 ? 1 : 0;
}

/* :5232-5287: record `index` (LINA_CHECKPOINT_LAST = -1 the last, -2 the one
 * before); *X malloc'd with *size = n * mesh_sizes[mesh_index - 1] values */
int lina_checkpoint_read(const char* path, const int* mesh_sizes, int n_meshes, int index, double** X, double* t,
                         int* size, int* mesh_index) {
    int n, m, i, mi = 0;
    FILE* f = fopen(path, "rb");
    *X = NULL;
    *size = 0;
    if (!f) {
        printf("could not open checkpoint file %s\n", path);
        return 0;
    }
    if (fread(&n, sizeof(int), 1, f) != 1 || fread(&m, sizeof(int), 1, f) != 1) {
        fclose(f);
        return 0;
    }
    if (index == LINA_CHECKPOINT_LAST) index = m - 1;
    if (index == LINA_CHECKPOINT_PREV_LAST) index = m - 2;
    if (index < 0 || index >= m) {
        printf("checkpoint read: index %d out of range 0-%d\n", index, m);
        fclose(f);
        return 0;
    }
    for (i = 0; i <= index; i++) {
        if (fread(t, sizeof(double), 1, f) != 1 || fread(&mi, sizeof(int), 1, f) != 1) {
            fclose(f);
            return 0;
        }
        if (mi < 1 || mi > n_meshes) {
            printf("checkpoint read: mesh index %d beyond the %d meshes\n", mi, n_meshes);
            fclose(f);
            return 0;
        }
        if (i < index) fseek(f, (long)n * mesh_sizes[mi - 1] * (long)sizeof(double), SEEK_CUR);
    }
    *size = n * mesh_sizes[mi - 1];
    *X = (double*)malloc((size_t)(*size > 0 ? *size : 1) * sizeof(double));
    if ((int)fread(*X, sizeof(double), (size_t)*size, f) != *size) {
        printf("checkpoint read: file %s ends inside record %d\n", path, index);
        free(*X);
        *X = NULL;
        fclose(f);
        return 0;
    }
    fclose(f);
    if (mesh_index) *mesh_index = mi;
    return 1;
}
