/* red_black_tree.h of the liblinalg_hip drop-in: the red-black tree of
 * LinA/include/red_black_tree.h with the same node and tree layouts and
 * signatures (fem_amd/host/rbtree.c).  Callers read the layout directly:
 * tree->root is a sentinel whose left child is the root of the tree, and
 * tree->nil is the sentinel every missing child points to. */
#ifndef LINALG_HIP_RB_TREE_H
#define LINALG_HIP_RB_TREE_H

#include "misc.h"
#include "stack.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rb_red_blk_node {
    void* key;
    void* info;
    int red; /* 0: black */
    struct rb_red_blk_node* left;
    struct rb_red_blk_node* right;
    struct rb_red_blk_node* parent;
} rb_red_blk_node;

/* Compare(a, b): 1 if *a > *b, -1 if *a < *b, 0 otherwise */
typedef struct rb_red_blk_tree {
    int (*Compare)(const void* a, const void* b);
    void (*DestroyKey)(void* a);
    void (*DestroyInfo)(void* a);
    void (*PrintKey)(const void* a);
    void (*PrintInfo)(void* a);
    rb_red_blk_node* root;
    rb_red_blk_node* nil;
} rb_red_blk_tree;

rb_red_blk_tree* RBTreeCreate(int (*CompFunc)(const void*, const void*), void (*DestFunc)(void*),
                              void (*InfoDestFunc)(void*), void (*PrintFunc)(const void*), void (*PrintInfo)(void*));
/* a key equal to stored ones goes after them in order; returns the new node */
rb_red_blk_node* RBTreeInsert(rb_red_blk_tree*, void* key, void* info);
/* PrintKey and PrintInfo of every node in key order */
void RBTreePrint(rb_red_blk_tree*);
int RBTreeCount(rb_red_blk_tree*);
/* removes the node, calling DestroyKey/DestroyInfo on its key and info;
 * every other node stays where the caller's pointers expect it */
void RBDelete(rb_red_blk_tree*, rb_red_blk_node*);
/* DestroyKey/DestroyInfo on every node, then all memory of the tree freed */
void RBTreeDestroy(rb_red_blk_tree*);
/* neighbours in key order; tree->nil when there is none */
rb_red_blk_node* TreePredecessor(rb_red_blk_tree*, rb_red_blk_node*);
rb_red_blk_node* TreeSuccessor(rb_red_blk_tree*, rb_red_blk_node*);
/* a node whose key compares equal, NULL when none does */
rb_red_blk_node* RBExactQuery(rb_red_blk_tree*, void*);
/* the last node in key order, NULL for an empty tree */
rb_red_blk_node* RBLargest(rb_red_blk_tree* tree);
/* the nodes with low <= key <= high, the smallest at the top of the stack */
stk_stack* RBEnumerate(rb_red_blk_tree* tree, void* low, void* high);
/* does nothing; a DestroyKey/DestroyInfo for trees that own nothing */
void NullFunction(void*);

#ifdef __cplusplus
}
#endif
#endif
