/*
 * fmm_oracle_tree.c -- CPU oracle restatement of the adaptive octree and the
 * interaction lists of d3tstrcr (src/d3tstrcr.f).  TEST INFRASTRUCTURE ONLY.
 *
 * Box numbering, point permutation and list contents follow the reference:
 *  - root cube = bounding box of sources and targets, side = largest extent,
 *    centre = midpoint (d3tallbem, d3tstrcr.f:1357-1395);
 *  - a box is split when it holds more than nbox sources or more than nbox
 *    targets (d3tstrcr.f:1490-1491), at most 100 levels (d3tstrcr.f:168-173);
 *  - children are produced in the order x fastest, then y, then z, points go
 *    to the upper half when strictly greater than the centre coordinate, the
 *    partition is stable (d3tsepa1/d3tsepa0, d3tstrcr.f:1705-1890) and empty
 *    children are dropped (d3tstrcr.f:1529);
 *  - boxes are numbered level by level (breadth first, d3tstrcr.f:1476-1576);
 *  - list 5 (colleagues) and list 2 come from the children of the parent's
 *    colleagues (d3tlsts, d3tstrcr.f:768-821), lists 1 and 3 from a descent
 *    into each colleague of a leaf (d3tlst31, d3tstrcr.f:912-1069), list 4 is
 *    the transpose of list 3 (d3tstrcr.f:887-901); "touching" is the corner
 *    test with tolerance side/10^4 (d3tifint, d3tstrcr.f:1075-1172).
 * Lists are stored CSR and sorted ascending (the reference keeps linked lists;
 * only summation order depends on the order).
 */
#include "fmm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int *v; int n, cap; } ivec;
static void ipush(ivec *a, int x) {
  if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 16; a->v = realloc(a->v, sizeof(int) * a->cap); }
  a->v[a->n++] = x;
}

/* d3tcentf / d3tcentc: centre of box (i,j,k) (1-based) at a level. */
static void box_center(const double *c0, double size, int level, const int *ijk, double *c) {
  double side = size / pow(2.0, level), side2 = side / 2;
  for (int d = 0; d < 3; d++) c[d] = (c0[d] - size / 2) + (ijk[d] - 1) * side + side2;
}

/* stable 8-way partition of idx[0..n) by comparison with the centre:
 * child = x + 2y + 4z with bit set when coord > centre. */
static void partition8(const double *pts, int *idx, int n, const double *c, int *cnt, int *tmp) {
  for (int k = 0; k < 8; k++) cnt[k] = 0;
  for (int i = 0; i < n; i++) {
    const double *x = pts + 3 * idx[i];
    int o = (x[0] > c[0]) | ((x[1] > c[1]) << 1) | ((x[2] > c[2]) << 2);
    cnt[o]++;
  }
  int off[8], s = 0;
  for (int k = 0; k < 8; k++) { off[k] = s; s += cnt[k]; }
  for (int i = 0; i < n; i++) {
    const double *x = pts + 3 * idx[i];
    int o = (x[0] > c[0]) | ((x[1] > c[1]) << 1) | ((x[2] > c[2]) << 2);
    tmp[off[o]++] = idx[i];
  }
  memcpy(idx, tmp, sizeof(int) * n);
}

/* d3tifint with corners from centre and side. */
static int touching(const orc_box *a, const orc_box *b, double size) {
  double sa = size / pow(2.0, a->level), sb = size / pow(2.0, b->level);
  double eps = (sa < sb ? sa : sb) / 10000;
  for (int d = 0; d < 3; d++) {
    double amin = a->center[d] - sa / 2, amax = amin + sa;
    double bmin = b->center[d] - sb / 2, bmax = bmin + sb;
    if (amin > bmax + eps || bmin > amax + eps) return 0;
  }
  return 1;
}

static int icmp(const void *a, const void *b) { return *(const int *)a - *(const int *)b; }

int orc_tree_build(const double *src, int ns, const double *trg, int nt, int nbox, orc_tree *t) {
  memset(t, 0, sizeof(*t));
  if (ns < 1) return 128;                     /* d3tstrcr.f:141-145 */
  double lo[3], hi[3];
  for (int d = 0; d < 3; d++) lo[d] = hi[d] = src[d];
  for (int i = 0; i < ns; i++)
    for (int d = 0; d < 3; d++) {
      if (src[3 * i + d] < lo[d]) lo[d] = src[3 * i + d];
      if (src[3 * i + d] > hi[d]) hi[d] = src[3 * i + d];
    }
  for (int i = 0; i < nt; i++)
    for (int d = 0; d < 3; d++) {
      if (trg[3 * i + d] < lo[d]) lo[d] = trg[3 * i + d];
      if (trg[3 * i + d] > hi[d]) hi[d] = trg[3 * i + d];
    }
  double size = hi[0] - lo[0];
  if (hi[1] - lo[1] > size) size = hi[1] - lo[1];
  if (hi[2] - lo[2] > size) size = hi[2] - lo[2];
  for (int d = 0; d < 3; d++) t->center[d] = (lo[d] + hi[d]) / 2;
  t->size = size;

  int cap = 1024, nb = 1;
  orc_box *B = calloc(cap, sizeof(orc_box));
  t->isrc = malloc(sizeof(int) * ns);
  t->itrg = malloc(sizeof(int) * (nt > 0 ? nt : 1));
  for (int i = 0; i < ns; i++) t->isrc[i] = i;
  for (int i = 0; i < nt; i++) t->itrg[i] = i;
  int *tmp = malloc(sizeof(int) * ((ns > nt ? ns : nt) + 1));
  B[0].level = 0; B[0].ijk[0] = B[0].ijk[1] = B[0].ijk[2] = 1; B[0].parent = -1;
  for (int k = 0; k < 8; k++) B[0].child[k] = -1;
  B[0].isrc = 0; B[0].nsrc = ns; B[0].itrg = 0; B[0].ntrg = nt;
  memcpy(B[0].center, t->center, sizeof(double) * 3);
  int ladd_cap = 128;
  t->ladd = malloc(sizeof(int) * ladd_cap);
  t->ladd[0] = 0; t->ladd[1] = 1;
  int nlev = 0;
  for (int level = 0; level < 100; level++) {
    int first = t->ladd[level], last = t->ladd[level + 1];
    int created = 0;
    for (int b = first; b < last; b++) {
      if (B[b].nsrc <= nbox && B[b].ntrg <= nbox) continue;
      int cs[8], ct[8];
      partition8(src, t->isrc + B[b].isrc, B[b].nsrc, B[b].center, cs, tmp);
      partition8(trg, t->itrg + B[b].itrg, B[b].ntrg, B[b].center, ct, tmp);
      int os = B[b].isrc, ot = B[b].itrg;
      for (int k = 0; k < 8; k++) {
        if (cs[k] == 0 && ct[k] == 0) { os += cs[k]; ot += ct[k]; continue; }
        if (nb == cap) { cap *= 2; B = realloc(B, sizeof(orc_box) * cap); }
        orc_box *c = &B[nb];
        memset(c, 0, sizeof(*c));
        c->level = level + 1;
        c->ijk[0] = 2 * (B[b].ijk[0] - 1) + 1 + (k & 1);
        c->ijk[1] = 2 * (B[b].ijk[1] - 1) + 1 + ((k >> 1) & 1);
        c->ijk[2] = 2 * (B[b].ijk[2] - 1) + 1 + ((k >> 2) & 1);
        c->parent = b;
        for (int q = 0; q < 8; q++) c->child[q] = -1;
        c->isrc = os; c->nsrc = cs[k]; c->itrg = ot; c->ntrg = ct[k];
        box_center(t->center, size, c->level, c->ijk, c->center);
        B[b].child[B[b].nchild++] = nb;
        os += cs[k]; ot += ct[k];
        nb++; created++;
      }
    }
    if (level + 2 >= ladd_cap) { ladd_cap *= 2; t->ladd = realloc(t->ladd, sizeof(int) * ladd_cap); }
    t->ladd[level + 2] = nb;
    if (created == 0) break;
    nlev = level + 1;
  }
  free(tmp);
  t->boxes = B; t->nboxes = nb; t->nlev = nlev;

  /* lists 5 and 2 */
  ivec L[5];
  ivec *per[5];
  memset(L, 0, sizeof(L));
  for (int k = 0; k < 5; k++) per[k] = calloc(nb, sizeof(ivec));
  for (int b = 1; b < nb; b++) {
    int dad = B[b].parent;
    int cand[1 + 1024], nc = 0;
    cand[nc++] = dad;
    for (int i = 0; i < per[4][dad].n; i++) cand[nc++] = per[4][dad].v[i];
    for (int i = 0; i < nc; i++) {
      const orc_box *cb = &B[cand[i]];
      for (int k = 0; k < cb->nchild; k++) {
        int kid = cb->child[k];
        if (kid == b) continue;
        if (touching(&B[kid], &B[b], size)) ipush(&per[4][b], kid);
        else ipush(&per[1][b], kid);
      }
    }
  }
  /* lists 1 and 3 (d3tlst31) */
  int *stack = malloc(sizeof(int) * 4096);
  for (int b = 0; b < nb; b++) {
    if (B[b].nchild > 0 || B[b].level == 0) continue;
    for (int i = 0; i < per[4][b].n; i++) {
      int sp = 0;
      stack[sp++] = per[4][b].v[i];
      while (sp > 0) {
        int j = stack[--sp];
        if (!touching(&B[b], &B[j], size)) { ipush(&per[2][b], j); continue; }
        if (B[j].nchild == 0) {
          ipush(&per[0][b], j);
          if (B[j].level != B[b].level) ipush(&per[0][j], b);
          continue;
        }
        for (int k = 0; k < B[j].nchild; k++) {
          if (sp >= 4096) { stack = realloc(stack, sizeof(int) * 2 * sp); }
          stack[sp++] = B[j].child[k];
        }
      }
    }
  }
  free(stack);
  for (int b = 0; b < nb; b++)
    for (int i = 0; i < per[2][b].n; i++) ipush(&per[3][per[2][b].v[i]], b);
  for (int k = 0; k < 5; k++) {
    t->lst_off[k] = malloc(sizeof(int) * (nb + 1));
    int tot = 0;
    for (int b = 0; b < nb; b++) tot += per[k][b].n;
    t->lst[k] = malloc(sizeof(int) * (tot > 0 ? tot : 1));
    int o = 0;
    for (int b = 0; b < nb; b++) {
      t->lst_off[k][b] = o;
      qsort(per[k][b].v, per[k][b].n, sizeof(int), icmp);
      if (per[k][b].n) memcpy(t->lst[k] + o, per[k][b].v, sizeof(int) * per[k][b].n);
      o += per[k][b].n;
      free(per[k][b].v);
    }
    t->lst_off[k][nb] = o;
    free(per[k]);
  }
  (void)L;
  return 0;
}

void orc_tree_free(orc_tree *t) {
  free(t->boxes); free(t->isrc); free(t->itrg); free(t->ladd);
  for (int k = 0; k < 5; k++) { free(t->lst_off[k]); free(t->lst[k]); }
  memset(t, 0, sizeof(*t));
}

/* Flattened view for tests: boxes20 in d3tgetb layout (1-based ids and
 * starts), permutations 1-based, lists per box ascending. */
int orc_tree_summary(const double *src, int ns, const double *trg, int nt, int nbox,
                     int *nboxes, int *nlev, int *boxes20, int maxboxes,
                     int *isrc, int *itrg, int *list_counts, int *lists, int maxlist) {
  orc_tree t;
  int ier = orc_tree_build(src, ns, trg, nt, nbox, &t);
  if (ier) return ier;
  *nboxes = t.nboxes; *nlev = t.nlev;
  if (t.nboxes > maxboxes) { orc_tree_free(&t); return 32; }
  for (int b = 0; b < t.nboxes; b++) {
    const orc_box *x = &t.boxes[b];
    int *o = boxes20 + 20 * b;
    memset(o, 0, sizeof(int) * 20);
    o[0] = x->level; o[1] = x->ijk[0]; o[2] = x->ijk[1]; o[3] = x->ijk[2];
    o[4] = x->parent + 1;
    for (int k = 0; k < x->nchild; k++) o[5 + k] = x->child[k] + 1;
    o[13] = x->isrc + 1; o[14] = x->nsrc; o[15] = x->itrg + 1; o[16] = x->ntrg;
    o[17] = x->nsrc == 0 ? 0 : (x->nchild ? 2 : 1);
    o[18] = x->ntrg == 0 ? 0 : (x->nchild ? 2 : 1);
  }
  for (int i = 0; i < ns; i++) isrc[i] = t.isrc[i] + 1;
  for (int i = 0; i < nt; i++) itrg[i] = t.itrg[i] + 1;
  int o = 0;
  for (int k = 0; k < 5; k++)
    for (int b = 0; b < t.nboxes; b++) {
      int a = t.lst_off[k][b], e = t.lst_off[k][b + 1];
      list_counts[k * maxboxes + b] = e - a;
      for (int i = a; i < e; i++) {
        if (o >= maxlist) { orc_tree_free(&t); return 32; }
        lists[o++] = t.lst[k][i] + 1;
      }
    }
  orc_tree_free(&t);
  return 0;
}
