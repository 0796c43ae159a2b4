/* Red-black tree, stack and helper functions of the LinA drop-in (host C),
 * behind the layouts of red_black_tree.h/stack.h that callers walk
 * themselves (FEM2D.c:881-949, K2_tree.c:76-85).
 *
 * The tree is the textbook one (Cormen et al., Introduction to Algorithms,
 * ch. 13) over two sentinels: tree->nil for every missing child and
 * tree->root, whose left child is the root.  Semantics kept from
 * LinA/red_black_tree.c:
 *   - a key that compares equal to stored keys goes to their right, i.e.
 *     after them in order (:174-202); the search compares Compare(node,key)
 *     and goes left only on 1;
 *   - the insertion and deletion rebalancing are the textbook cases, so a
 *     sequence of operations leaves the same shape and colours (callers
 *     descend from tree->root->left with their own, tolerance-based
 *     comparisons, so the shape is observable);
 *   - RBDelete(z) frees z only: when z has two children its successor node
 *     takes z's place, so pointers to every other node stay valid (:620-…);
 *   - RBLargest and RBExactQuery return NULL (not tree->nil) when nothing is
 *     found (:513-523, :486-501); TreePredecessor/TreeSuccessor return
 *     tree->nil;
 *   - RBEnumerate stacks the nodes of [low, high] from the largest down, so
 *     the smallest is on top (:683-…).
 * Mirror-image cases are written once over a side index d (0 left, 1 right). */
#include "dropin.h"

void Assert(int assertion, char* error) { /* misc.c */
    if (!assertion) {
        printf("Assertion Failed: %s\n", error);
        exit(-1);
    }
}

void* SafeMalloc(size_t size) { /* misc.c */
    void* p = malloc(size ? size : 1);
    if (!p) {
        printf("memory overflow: malloc failed in SafeMalloc.  Exiting Program.\n");
        exit(-1);
    }
    return p;
}

void NullFunction(void* junk) { (void)junk; }

/* ---------------------------------------------------------------- stack */

stk_stack* StackCreate() {
    stk_stack* s = (stk_stack*)SafeMalloc(sizeof(stk_stack));
    s->top = s->tail = NULL;
    return s;
}

void StackPush(stk_stack* theStack, DATA_TYPE newInfoPointer) {
    stk_stack_node* nd = (stk_stack_node*)SafeMalloc(sizeof(stk_stack_node));
    nd->info = newInfoPointer;
    nd->next = theStack->top;
    if (!theStack->top) theStack->tail = nd;
    theStack->top = nd;
}

void* StackPop(stk_stack* theStack) {
    stk_stack_node* nd = theStack->top;
    void* info;
    if (!nd) return NULL;
    info = nd->info;
    theStack->top = nd->next;
    if (!theStack->top) theStack->tail = NULL;
    free(nd);
    return info;
}

int StackNotEmpty(stk_stack* theStack) { return theStack != NULL && theStack->top != NULL; }

stk_stack* StackJoin(stk_stack* stack1, stk_stack* stack2) {
    if (!stack1->tail) {
        free(stack1);
        return stack2;
    }
    stack1->tail->next = stack2->top;
    if (stack2->tail) stack1->tail = stack2->tail;
    free(stack2);
    return stack1;
}

void StackDestroy(stk_stack* theStack, void DestFunc(void* a)) {
    stk_stack_node* nd;
    if (!theStack) return;
    nd = theStack->top;
    while (nd) {
        stk_stack_node* next = nd->next;
        DestFunc(nd->info);
        free(nd);
        nd = next;
    }
    free(theStack);
}

/* ----------------------------------------------------------------- tree */

static rb_red_blk_node** link_of(rb_red_blk_node* p, int d) { return d ? &p->right : &p->left; }

/* the side of its parent x hangs on */
static int side_of(const rb_red_blk_node* x) { return x == x->parent->right; }

/* x's child on side d moves up into x's place; x becomes its child on the
 * other side (d = 1: left rotation, d = 0: right rotation) */
static void rotate_up(rb_red_blk_tree* t, rb_red_blk_node* x, int d) {
    rb_red_blk_node* y = *link_of(x, d);
    rb_red_blk_node* inner = *link_of(y, !d);
    *link_of(x, d) = inner;
    if (inner != t->nil) inner->parent = x;
    y->parent = x->parent;
    *link_of(x->parent, side_of(x)) = y;
    *link_of(y, !d) = x;
    x->parent = y;
}

rb_red_blk_tree* RBTreeCreate(int (*CompFunc)(const void*, const void*), void (*DestFunc)(void*),
                              void (*InfoDestFunc)(void*), void (*PrintFunc)(const void*), void (*PrintInfo)(void*)) {
    rb_red_blk_tree* t = (rb_red_blk_tree*)SafeMalloc(sizeof(rb_red_blk_tree));
    rb_red_blk_node* nil = (rb_red_blk_node*)SafeMalloc(sizeof(rb_red_blk_node));
    rb_red_blk_node* root = (rb_red_blk_node*)SafeMalloc(sizeof(rb_red_blk_node));
    t->Compare = CompFunc;
    t->DestroyKey = DestFunc;
    t->DestroyInfo = InfoDestFunc;
    t->PrintKey = PrintFunc;
    t->PrintInfo = PrintInfo;
    nil->key = nil->info = NULL;
    nil->red = 0;
    nil->left = nil->right = nil->parent = nil;
    root->key = root->info = NULL;
    root->red = 0;
    root->left = root->right = root->parent = nil;
    t->nil = nil;
    t->root = root;
    return t;
}

rb_red_blk_node* RBTreeInsert(rb_red_blk_tree* tree, void* key, void* info) {
    rb_red_blk_node* const nil = tree->nil;
    rb_red_blk_node* z = (rb_red_blk_node*)SafeMalloc(sizeof(rb_red_blk_node));
    rb_red_blk_node* p = tree->root;
    rb_red_blk_node* x = tree->root->left;
    int d = 0;
    z->key = key;
    z->info = info;
    z->left = z->right = nil;
    z->red = 1;
    while (x != nil) {
        p = x;
        d = tree->Compare(x->key, key) != 1; /* right unless the node's key is greater */
        x = *link_of(x, d);
    }
    z->parent = p;
    *link_of(p, p == tree->root ? 0 : d) = z;
    /* restore: no red node with a red parent (the root sentinel is black) */
    x = z;
    while (x->parent->red) {
        rb_red_blk_node* g = x->parent->parent;
        const int pd = side_of(x->parent);
        rb_red_blk_node* uncle = *link_of(g, !pd);
        if (uncle->red) {
            x->parent->red = 0;
            uncle->red = 0;
            g->red = 1;
            x = g;
        } else {
            if (side_of(x) != pd) { /* inner grandchild: make it outer */
                x = x->parent;
                rotate_up(tree, x, !pd);
            }
            x->parent->red = 0;
            x->parent->parent->red = 1;
            rotate_up(tree, x->parent->parent, pd);
        }
    }
    tree->root->left->red = 0;
    return z;
}

/* the leftmost (d = 0) or rightmost (d = 1) node below x (x != nil) */
static rb_red_blk_node* extreme(rb_red_blk_tree* t, rb_red_blk_node* x, int d) {
    while (*link_of(x, d) != t->nil) x = *link_of(x, d);
    return x;
}

/* the neighbour of x on side d in key order, tree->nil when there is none */
static rb_red_blk_node* neighbour(rb_red_blk_tree* t, rb_red_blk_node* x, int d) {
    if (*link_of(x, d) != t->nil) return extreme(t, *link_of(x, d), !d);
    while (x->parent != t->root && side_of(x) == d) x = x->parent;
    return x->parent == t->root ? t->nil : x->parent;
}

rb_red_blk_node* TreeSuccessor(rb_red_blk_tree* tree, rb_red_blk_node* x) { return neighbour(tree, x, 1); }

rb_red_blk_node* TreePredecessor(rb_red_blk_tree* tree, rb_red_blk_node* x) { return neighbour(tree, x, 0); }

static void count_below(rb_red_blk_tree* t, rb_red_blk_node* x, int* c) {
    while (x != t->nil) {
        (*c)++;
        count_below(t, x->left, c);
        x = x->right;
    }
}

int RBTreeCount(rb_red_blk_tree* tree) {
    int c = 0;
    count_below(tree, tree->root->left, &c);
    return c;
}

static void print_below(rb_red_blk_tree* t, rb_red_blk_node* x) {
    if (x == t->nil) return;
    print_below(t, x->left);
    t->PrintKey(x->key);
    t->PrintInfo(x->info);
    print_below(t, x->right);
}

void RBTreePrint(rb_red_blk_tree* tree) { print_below(tree, tree->root->left); }

static void destroy_below(rb_red_blk_tree* t, rb_red_blk_node* x) {
    if (x == t->nil) return;
    destroy_below(t, x->left);
    destroy_below(t, x->right);
    t->DestroyKey(x->key);
    t->DestroyInfo(x->info);
    free(x);
}

void RBTreeDestroy(rb_red_blk_tree* tree) {
    destroy_below(tree, tree->root->left);
    free(tree->root);
    free(tree->nil);
    free(tree);
}

rb_red_blk_node* RBExactQuery(rb_red_blk_tree* tree, void* q) {
    rb_red_blk_node* x = tree->root->left;
    while (x != tree->nil) {
        const int c = tree->Compare(x->key, q);
        if (c == 0) return x;
        x = c == 1 ? x->left : x->right;
    }
    return NULL;
}

rb_red_blk_node* RBLargest(rb_red_blk_tree* tree) {
    rb_red_blk_node* x = tree->root->left;
    return x == tree->nil ? NULL : extreme(tree, x, 1);
}

/* v takes u's place under u's parent (v may be nil: its parent is set too) */
static void replace_in_parent(rb_red_blk_tree* t, rb_red_blk_node* u, rb_red_blk_node* v) {
    (void)t;
    *link_of(u->parent, side_of(u)) = v;
    v->parent = u->parent;
}

/* x carries an extra black after a black node left its path */
static void delete_fixup(rb_red_blk_tree* t, rb_red_blk_node* x) {
    while (!x->red && x != t->root->left) {
        const int d = side_of(x);
        rb_red_blk_node* w = *link_of(x->parent, !d);
        if (w->red) {
            w->red = 0;
            x->parent->red = 1;
            rotate_up(t, x->parent, !d);
            w = *link_of(x->parent, !d);
        }
        if (!w->left->red && !w->right->red) {
            w->red = 1;
            x = x->parent;
        } else {
            if (!(*link_of(w, !d))->red) { /* near child red, far child black */
                (*link_of(w, d))->red = 0;
                w->red = 1;
                rotate_up(t, w, d);
                w = *link_of(x->parent, !d);
            }
            w->red = x->parent->red;
            x->parent->red = 0;
            (*link_of(w, !d))->red = 0;
            rotate_up(t, x->parent, !d);
            x = t->root->left;
        }
    }
    x->red = 0;
}

void RBDelete(rb_red_blk_tree* tree, rb_red_blk_node* z) {
    rb_red_blk_node* const nil = tree->nil;
    rb_red_blk_node* x;
    int removed_black = !z->red;
    if (z->left == nil || z->right == nil) {
        x = z->left == nil ? z->right : z->left;
        replace_in_parent(tree, z, x);
    } else {
        rb_red_blk_node* y = extreme(tree, z->right, 0); /* z's successor moves into z's place */
        removed_black = !y->red;
        x = y->right;
        if (y->parent == z) {
            x->parent = y;
        } else {
            replace_in_parent(tree, y, x);
            y->right = z->right;
            y->right->parent = y;
        }
        replace_in_parent(tree, z, y);
        y->left = z->left;
        y->left->parent = y;
        y->red = z->red;
    }
    if (removed_black) delete_fixup(tree, x);
    tree->DestroyKey(z->key);
    tree->DestroyInfo(z->info);
    free(z);
}

stk_stack* RBEnumerate(rb_red_blk_tree* tree, void* low, void* high) {
    stk_stack* s = StackCreate();
    rb_red_blk_node* x = tree->root->left;
    rb_red_blk_node* best = tree->nil; /* the last node in order with key <= high */
    while (x != tree->nil) {
        if (tree->Compare(x->key, high) == 1) {
            x = x->left;
        } else {
            best = x;
            x = x->right;
        }
    }
    while (best != tree->nil && tree->Compare(low, best->key) != 1) {
        StackPush(s, best);
        best = TreePredecessor(tree, best);
    }
    return s;
}
