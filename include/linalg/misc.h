/* misc.h of the liblinalg_hip drop-in: the assertion and allocation helpers
 * of LinA/include/misc.h, same signatures (fem_amd/host/rbtree.c). */
#ifndef LINALG_HIP_MISC_H
#define LINALG_HIP_MISC_H

#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* prints "Assertion Failed: <error>" and exits with -1 when assertion is 0 */
void Assert(int assertion, char* error);
/* malloc that prints a message and exits with -1 when it fails */
void* SafeMalloc(size_t size);

#ifdef __cplusplus
}
#endif
#endif
