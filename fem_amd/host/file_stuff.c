/* File helpers of the LinA drop-in (host C): the functions of
 * LinA/File_stuff.c that Crack and FEM2D call to read their parameter and
 * list files.  Results on the inputs the reference handles are the
 * reference's; deliberate differences:
 *   - a file that cannot be opened gives NULL and a message instead of
 *     exit(0) (open_file) or a crash (write_lists, read_lists);
 *   - split keeps tokens of any length (each buffer holds at least 256
 *     bytes, as callers may write into them) and the whole string (the
 *     reference truncates at 2048 characters and 256 per token without a
 *     terminator);
 *   - read_lists reads lines of any length (the reference stops at 511
 *     characters per line);
 *   - a byte 0xFF inside a line is data, not end of file. */
#include <string.h>

#include "dropin.h"

FILE* open_file(char* dir, char* name, char* mode) { /* :6-22 */
    FILE* f;
    char* path;
    if (name[0] == '/') {
        path = strdup(name);
    } else {
        path = (char*)malloc(strlen(dir) + strlen(name) + 2);
        sprintf(path, "%s/%s", dir, name);
    }
    f = fopen(path, mode);
    if (!f) {
        if (strcmp(mode, "r") == 0) printf("file %s does not exist\n", path);
        else printf("file %s could not be created or opened\n", path);
    }
    free(path);
    return f;
}

FILE* Open_file(char* filename, char* mode) { /* :24-30 */
    FILE* f = fopen(filename, mode);
    if (!f && strcmp(mode, "r") == 0) printf("file %s does not exist\n", filename);
    return f;
}

void append_data(FILE* file, double* data, int size, int var_num, double time) { /* :32-48 */
    int i, j;
    fprintf(file, "t = %f\n", time);
    for (i = 0; i < size; i++) {
        for (j = 0; j < var_num; j++) fprintf(file, "%lf ", data[(size_t)size * j + i]);
        fputc('\n', file);
    }
    fflush(file);
}

char** split(char* String, char* delimiter, int* Number) { /* :95-116 */
    char** parts = NULL;
    const char* s = String;
    *Number = 0;
    if (!String) return NULL;
    for (;;) {
        size_t len;
        char* tok;
        s += strspn(s, delimiter);
        if (!*s) break;
        len = strcspn(s, delimiter);
        tok = (char*)calloc(len + 1 > 256 ? len + 1 : 256, 1);
        memcpy(tok, s, len);
        parts = (char**)realloc(parts, (size_t)(*Number + 1) * sizeof(char*));
        parts[(*Number)++] = tok;
        s += len;
    }
    return parts;
}

/* characters up to '\n' or end of file into buffer (at most cap - 1 of
 * them): 1 after '\n', 0 at end of file, -1 when the line does not fit */
static int read_into(FILE* file, char* buffer, int cap) {
    int c, i = 0;
    for (;;) {
        c = fgetc(file);
        if (c == '\n' || c == EOF) break;
        buffer[i++] = (char)c;
        if (i == cap - 1) {
            buffer[i] = '\0';
            printf("line too long: %s\n", buffer);
            return -1;
        }
    }
    buffer[i] = '\0';
    return c == EOF ? 0 : 1;
}

int readline(FILE* file, char* buffer) { return read_into(file, buffer, 512); } /* :143-162 */

int readLine(FILE* file, char* buffer, int buff_size) { return read_into(file, buffer, buff_size); } /* :164-182 */

void write_lists(char* path, char* name, int** Lists, int number, int size) { /* :184-206 */
    int i, j;
    char* fn = (char*)malloc(strlen(path) + strlen(name) + 2);
    FILE* f;
    sprintf(fn, "%s/%s", path, name);
    f = fopen(fn, "w");
    if (!f) {
        printf("file %s could not be created\n", fn);
        free(fn);
        return;
    }
    for (i = 0; i < size; i++) {
        fprintf(f, "%d", Lists[0][i]);
        for (j = 1; j < number; j++) fprintf(f, " %d", Lists[j][i]);
        fputc('\n', f);
    }
    fclose(f);
    free(fn);
}

/* one whole line into *buf (grown as needed): 1 after '\n', 0 at end of file */
static int read_whole(FILE* file, char** buf, size_t* cap) {
    size_t i = 0;
    int c;
    for (;;) {
        c = fgetc(file);
        if (c == '\n' || c == EOF) break;
        if (i + 1 >= *cap) {
            *cap = *cap ? 2 * *cap : 256;
            *buf = (char*)realloc(*buf, *cap);
        }
        (*buf)[i++] = (char)c;
    }
    if (!*buf) {
        *cap = 256;
        *buf = (char*)malloc(*cap);
    }
    (*buf)[i] = '\0';
    return c == EOF ? 0 : 1;
}

/* :208-241: one list per column of the first line, each line a further entry;
 * stops at the first empty line or at the last line that ends with '\n';
 * *Number is the token count of the last line read, as in the reference */
int** read_lists(char* path, char* name, int* Number, int* Size) {
    int i, cols = 0;
    int** res = NULL;
    char* line = NULL;
    size_t cap = 0;
    char* fn = (char*)malloc(strlen(path) + strlen(name) + 2);
    FILE* f;
    sprintf(fn, "%s/%s", path, name);
    f = fopen(fn, "r");
    *Size = 0;
    *Number = 0;
    if (!f) {
        printf("file %s does not exist\n", fn);
        free(fn);
        return NULL;
    }
    free(fn);
    while (read_whole(f, &line, &cap)) {
        char** parts = split(line, " ", Number);
        if (*Number <= 0) break;
        if (*Size == 0) {
            cols = *Number;
            res = (int**)calloc((size_t)cols, sizeof(int*));
        }
        (*Size)++;
        for (i = 0; i < *Number; i++) {
            if (i < cols) { /* columns beyond the first line's are dropped */
                res[i] = (int*)realloc(res[i], (size_t)*Size * sizeof(int));
                res[i][*Size - 1] = atoi(parts[i]);
            }
            free(parts[i]);
        }
        free(parts);
    }
    free(line);
    fclose(f);
    return res;
}
