/* File_stuff.h of the liblinalg_hip drop-in: the file helpers of
 * LinA/include/File_stuff.h, same signatures (fem_amd/host/file_stuff.c). */
#ifndef LINALG_HIP_FILE_STUFF_H
#define LINALG_HIP_FILE_STUFF_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

/* one line (without its '\n') into buffer: 1 when a '\n' ended it, 0 at end
 * of file, -1 when it would not fit (buffer of 512 bytes for readline,
 * buff_size for readLine) */
int readline(FILE* file, char* buffer);
int readLine(FILE* file, char* buffer, int buff_size);
/* the tokens of String separated by any character of delimiter, each in its
 * own malloc'd buffer of at least 256 bytes, *Number of them; NULL when
 * there is none */
char** split(char* String, char* delimiter, int* Number);
/* dir/name (name alone when it starts with '/'); NULL and a message when it
 * cannot be opened */
FILE* open_file(char* dir, char* name, char* mode);
FILE* Open_file(char* filename, char* mode);
/* "t = <time>" then one line per entry i of the var_num fields data[size j + i] */
void append_data(FILE* file, double* data, int size, int var_num, double time);
/* number lists of size integers written column-wise, one line per index */
void write_lists(char* path, char* name, int** Lists, int number, int size);
int** read_lists(char* path, char* name, int* Number, int* Size);

#ifdef __cplusplus
}
#endif
#endif
