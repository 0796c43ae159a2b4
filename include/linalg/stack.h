/* stack.h of the liblinalg_hip drop-in: the singly linked stack of
 * LinA/include/stack.h with the same node and stack layouts (callers walk
 * top -> next ... tail themselves) and signatures (fem_amd/host/rbtree.c). */
#ifndef LINALG_HIP_STACK_H
#define LINALG_HIP_STACK_H

#include "misc.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef DATA_TYPE
#define DATA_TYPE void*
#endif

typedef struct stk_stack_node {
    DATA_TYPE info;
    struct stk_stack_node* next;
} stk_stack_node;

/* top: the last pushed node; tail: the first pushed (NULL when empty) */
typedef struct stk_stack {
    stk_stack_node* top;
    stk_stack_node* tail;
} stk_stack;

/* stack1's nodes followed by stack2's; frees one of the two headers and
 * returns the other */
stk_stack* StackJoin(stk_stack* stack1, stk_stack* stack2);
stk_stack* StackCreate();
void StackPush(stk_stack* theStack, DATA_TYPE newInfoPointer);
/* NULL when empty */
void* StackPop(stk_stack* theStack);
/* nonzero when theStack holds a node */
int StackNotEmpty(stk_stack*);
/* DestFunc on every info from top to tail, then the nodes and the header freed */
void StackDestroy(stk_stack* theStack, void DestFunc(void* a));

#ifdef __cplusplus
}
#endif
#endif
