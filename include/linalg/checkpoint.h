/* checkpoint.h of the liblinalg_hip drop-in: Crack's field checkpoint file
 * data.bin (fem_amd/host/checkpoint.c).  No counterpart in LinA: the layout
 * of Crack's init_binary_save / binary_save / binary_load
 * (Crack/crack.c:5095-5287) -- int n (fields per node), int m (records),
 * then m records of double t, int mesh_index (1-based) and
 * n * mesh_size[mesh_index - 1] doubles, machine byte order, no padding --
 * so that runs of this library and Crack runs read each other's fields.
 * Return SUCCESS (1) or FAILED (0) like Crack's functions. */
#ifndef LINALG_HIP_CHECKPOINT_H
#define LINALG_HIP_CHECKPOINT_H

#ifdef __cplusplus
extern "C" {
#endif

#define LINA_CHECKPOINT_LAST (-1)      /* LAST_INDEX, crack.c:48 */
#define LINA_CHECKPOINT_PREV_LAST (-2) /* PREV_LAST_INDEX, crack.c:49 */

/* a new file with n_fields and no record; names (NULL or n_fields strings)
 * go to "<path>.info", the count on the first line, one name a line */
int lina_checkpoint_init(const char* path, int n_fields, const char* const* names);
/* binary_save: one record (size = n_fields * mesh size values) at the end */
int lina_checkpoint_append(const char* path, int n_fields, const double* X, int size, double t, int mesh_index);
/* the header: fields per node and records */
int lina_checkpoint_info(const char* path, int* n_fields, int* records);
/* load_mesh_sizes: the first token of <dir>/meshes/<mesh_base>.<k>.node for
 * k = 1, 2, ... while the file exists; *sizes malloc'd */
int lina_checkpoint_mesh_sizes(const char* dir, const char* mesh_base, int** sizes, int* count);
/* binary_load: record `index` (or LINA_CHECKPOINT_LAST / _PREV_LAST); *X
 * malloc'd with *size values, *t its time, *mesh_index its mesh */
int lina_checkpoint_read(const char* path, const int* mesh_sizes, int n_meshes, int index, double** X, double* t,
                         int* size, int* mesh_index);

#ifdef __cplusplus
}
#endif
#endif
