/*
 * Precision-independent part of the QLL C API served by libqll_amd:
 * aligned allocation, the host site layout (layoutSetup/Index/Coord/Shift/
 * Subset), the team sum gsum, and the registry that ties QLL structs to their
 * device objects.
 */
#define _POSIX_C_SOURCE 200112L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define QLL_NO_GENERIC_NAMES
#include "common.h"
#include "qll_amd.h"
#include "qll_host.h"

int myrank = 0;
int qllVlen[2] = {1, 1};   /* [0] float API, [1] double API */

void qllSetVlen(int vlenF, int vlenD)
{
  if (vlenF >= 1) qllVlen[0] = vlenF;
  if (vlenD >= 1) qllVlen[1] = vlenD;
}

/* lib/common.c:9-28 */
void *myalloc(size_t size)
{
  void *b = NULL;
  if (size == 0) size = 64;
  int err = posix_memalign(&b, 64, size);
  if (err) {
    printf("posix_memalign(&%p, %li, %li) failed!\n", b, (long)64, (long)size);
    exit(1);
  }
  return b;
}

/* team sum of n doubles (lib/common.c:30-72 on one rank): every thread of
 * the team ends with the total; thread 0 accumulates in thread order */
void gsum(double *x, int n, ThreadCtx *TC)
{
  if (!TC || TC->nid <= 1) return;
  int tid = TC->tid;
  TC->sum[tid] = x;
  OMP_BARRIER;
  if (tid == 0) {
    for (int t = 1; t < TC->nid; t++)
      for (int i = 0; i < n; i++) x[i] += TC->sum[t][i];
  }
  OMP_BARRIER;
  if (tid != 0)
    for (int i = 0; i < n; i++) x[i] = TC->sum[0][i];
  OMP_BARRIER;
}

/* ------------------------------------------------------------- layout */
/* layoutSubset (lib/layout.c:5-19) */
void layoutSubset(Subset *s, Layout *l, char *sub)
{
  s->begin = 0;
  s->end = l->nSites;
  s->beginOuter = 0;
  s->endOuter = l->nSitesOuter;
  if (sub[0] == 'e') {
    s->end = l->nEven;
    s->endOuter = l->nEvenOuter;
  } else if (sub[0] == 'o') {
    s->begin = l->nOdd;
    s->beginOuter = l->nOddOuter;
  }
}

/* layoutSetup (lib/layout.c:21-80): outer geometry, checkerboard fix-up of
 * the inner lanes, even/odd counts.  Exits like QLL on unusable geometry. */
void layoutSetup(Layout *l)
{
  int nd = l->nDim;
  l->outerGeom = myalloc(nd * sizeof(int));
  l->localGeom = myalloc(nd * sizeof(int));
  int pvol = 1, lvol = 1, ovol = 1, icb = 0, icbd = -1;
  for (int i = 0; i < nd; i++) {
    l->localGeom[i] = l->physGeom[i] / l->rankGeom[i];
    l->outerGeom[i] = l->localGeom[i] / l->innerGeom[i];
    pvol *= l->physGeom[i];
    lvol *= l->localGeom[i];
    ovol *= l->outerGeom[i];
    if (l->innerGeom[i] > 1 && (l->outerGeom[i] & 1) == 1) icb++;
    if (l->innerGeom[i] == 1 && (l->outerGeom[i] & 1) == 0) icbd = i;
  }
  if (icb == 0) {
    icbd = 0;
  } else {
    if (icbd < 0) {
      if (l->myrank == 0) printf("not enough 2's in localGeom\n");
      exit(-1);
    }
    icb = l->outerGeom[icbd] / 2;
    if ((icb & 1) == 0) {
      if (l->myrank == 0) printf("error in cb choice\n");
      exit(-1);
    }
  }
  l->physVol = pvol;
  l->nSites = lvol;
  l->nOdd = lvol / 2;
  l->nEven = lvol - l->nOdd;
  l->nSitesOuter = ovol;
  l->nOddOuter = ovol / 2;
  l->nEvenOuter = ovol - l->nOddOuter;
  l->nSitesInner = l->nSites / l->nSitesOuter;
  l->innerCb = icb;
  l->innerCbDir = icbd;
}

/* x[0] fastest; with d: coordinate divided by d first (lib/layout.c:92-103) */
static int lexi(const int *x, const int *s, const int *d, int nd)
{
  int r = 0;
  for (int i = nd - 1; i >= 0; i--) {
    int xx = x[i];
    if (d) xx /= d[i];
    r = r * s[i] + (xx % s[i]);
  }
  return r;
}

/* layoutIndex (lib/layout.c:120-142) */
void layoutIndex(Layout *l, LayoutIndex *li, int coords[])
{
  int nd = l->nDim;
  int x[8];
  for (int i = 0; i < nd; i++) x[i] = coords[i];
  int ri = lexi(x, l->rankGeom, l->localGeom, nd);
  int ii = lexi(x, l->innerGeom, l->outerGeom, nd);
  int ib = 0;
  for (int i = 0; i < nd; i++) ib += ((x[i] / l->outerGeom[i]) % l->innerGeom[i]) * l->outerGeom[i];
  ib &= 1;
  x[l->innerCbDir] += l->innerCb * ib;
  int oi = lexi(x, l->outerGeom, NULL, nd);
  int p = 0;
  for (int i = 0; i < nd; i++) p += coords[i];
  int oi2 = (p & 1) ? (oi + l->nSitesOuter) / 2 : oi / 2;
  li->rank = ri;
  li->index = oi2 * l->nSitesInner + ii;
}

/* layoutCoord (lib/layout.c:144-205), including QLL's round-trip check */
void layoutCoord(Layout *l, int *coords, LayoutIndex *li)
{
  int nd = l->nDim, cr[8], p = 0;
  int r = li->rank;
  for (int i = 0; i < nd; i++) {
    cr[i] = r % l->rankGeom[i];
    r /= l->rankGeom[i];
  }
  int ll = li->index % l->nSitesInner, ib = 0;
  for (int i = 0; i < nd; i++) {
    int w = l->innerGeom[i], wl = l->outerGeom[i], k = ll % w;
    int c = l->localGeom[i] * cr[i] + k * wl;
    cr[i] = c;
    p += c;
    ll /= w;
    ib += k * wl;
  }
  ib &= 1;
  int ii = li->index / l->nSitesInner;
  if (ii >= l->nEvenOuter) {
    ii -= l->nEvenOuter;
    p++;
  }
  ii *= 2;
  for (int i = 0; i < nd; i++) {
    int wl = l->outerGeom[i], k = ii % wl;
    if (i == l->innerCbDir) k = (k + l->innerCb * ib) % wl;
    coords[i] = k;
    p += k;
    ii /= wl;
  }
  if (p & 1) {
    for (int i = 0; i < nd; i++) {
      int wl = l->outerGeom[i];
      if (i == l->innerCbDir) coords[i] = (coords[i] + l->innerCb * ib) % wl;
      coords[i]++;
      if (coords[i] >= wl) {
        coords[i] = 0;
        if (i == l->innerCbDir) coords[i] = (coords[i] + l->innerCb * ib) % wl;
      } else {
        if (i == l->innerCbDir) coords[i] = (coords[i] + l->innerCb * ib) % wl;
        break;
      }
    }
  }
  for (int i = 0; i < nd; i++) coords[i] += cr[i];
  LayoutIndex li2;
  layoutIndex(l, &li2, coords);
  if (li->rank != li2.rank || li->index != li2.index) {
    printf("error: bad coord: %i,%i -> %i,%i\n", li->rank, li->index, li2.rank, li2.index);
    exit(-1);
  }
}

/* layoutShift (lib/layout.c:207-217) */
void layoutShift(Layout *l, LayoutIndex *li, LayoutIndex *li2, int *disp)
{
  int nd = l->nDim, x[8];
  layoutCoord(l, x, li2);
  for (int i = 0; i < nd; i++) x[i] = (x[i] + disp[i] + l->physGeom[i]) % l->physGeom[i];
  layoutIndex(l, li, x);
}

/* makeShiftMultiSub (lib/shifts.c:516-736) for one rank: si[n] describes
 * dst(x) = src(x - disp[n]) on the outer sites of subs[n].  libqll_amd's
 * shifts and transporters only read disp (the device computes neighbours
 * from coordinates); the site tables are filled for callers that inspect
 * them, as QLL fills them when nothing crosses a rank: sidx[i] = the outer
 * site holding the source of lane 0 of outer site i, pidx[i] the same or
 * -(that)-2 when the lanes arrive permuted (perm = the source lane of lane
 * 0), -1 outside the subset.  No message member is set. */
void makeShiftMultiSub(ShiftIndices *si[], Layout *l, int *disp[], char *subs[], int ndisp)
{
  int nd = l->nDim, vv = l->nSitesOuter, NI = l->nSitesInner;
  int perm = 0;
  for (int n = 0; n < ndisp; n++) {
    ShiftIndices *s = si[n];
    memset(s, 0, sizeof *s);
    s->disp = myalloc(nd * sizeof(int));
    memcpy(s->disp, disp[n], nd * sizeof(int));
    s->pidx = myalloc(vv * sizeof(int));
    s->sidx = myalloc(vv * sizeof(int));
    s->sendSites = myalloc(vv * sizeof(int));
    s->vv = vv;
    Subset sub;
    layoutSubset(&sub, l, subs[n]);
    int back[8];
    for (int k = 0; k < nd; k++) back[k] = ((-disp[n][k]) % l->physGeom[k] + l->physGeom[k]) % l->physGeom[k];
    for (int i = 0; i < vv; i++) {
      s->pidx[i] = s->sidx[i] = -1;
      if (i < sub.beginOuter || i >= sub.endOuter) continue;
      LayoutIndex dst = {0, i * NI}, src;
      layoutShift(l, &src, &dst, back);
      int so = src.index / NI, lane = src.index % NI;
      s->sidx[i] = so;
      s->pidx[i] = so;
      if (lane != 0) {
        perm = lane;
        s->pidx[i] = -so - 2;
      }
    }
  }
  for (int n = 0; n < ndisp; n++) si[n]->perm = perm;
}

void makeShiftMulti(ShiftIndices *si[], Layout *l, int *disp[], int ndisp)
{
  char **subs = malloc((ndisp > 0 ? ndisp : 1) * sizeof(char *));
  for (int k = 0; k < ndisp; k++) subs[k] = "all";
  makeShiftMultiSub(si, l, disp, subs, ndisp);
  free(subs);
}

void makeShiftSub(ShiftIndices *si, Layout *l, int *disp, char *sub)
{
  makeShiftMultiSub(&si, l, &disp, &sub, 1);
}

void makeShift(ShiftIndices *si, Layout *l, int *disp) { makeShiftSub(si, l, disp, "all"); }

void freeShift(ShiftIndices *si)
{
  free(si->disp);
  free(si->pidx);
  free(si->sidx);
  free(si->sendSites);
  si->disp = si->pidx = si->sidx = si->sendSites = NULL;
}

/* partitionGeom (lib/stagDslash.c:751-786) */
void qllPartitionGeom(int *lx, int *nx, const int *x, int n, int nd, int dist)
{
  for (int i = 0; i < nd; i++) {
    nx[i] = 1;
    lx[i] = x[i];
  }
  int ww = n;
  while (ww > 1) {
    int k = nd - 1;
    while (lx[k] & 1) {
      k--;
      if (k < 0) {
        printf("not enough 2's in partitioned geom:");
        for (int i = 0; i < nd; i++) printf(" %i", x[i]);
        printf(" /%i\n", n);
        exit(-1);
      }
    }
    for (int i = k - 1; i >= 0; i--) {
      if ((lx[i] & 1) == 0) {
        if (dist == 0) {
          if (lx[i] > lx[k] || (lx[i] == lx[k] && nx[i] > nx[k])) k = i;
        } else {
          if (nx[i] < nx[k] || (nx[i] == nx[k] && lx[i] > lx[k])) k = i;
        }
      }
    }
    nx[k] *= 2;
    lx[k] /= 2;
    ww /= 2;
  }
}

void qllToDevLayout(const Layout *l, qamd_qll_layout *q)
{
  memset(q, 0, sizeof *q);
  for (int i = 0; i < 4; i++) {
    q->physGeom[i] = l->physGeom[i];
    q->innerGeom[i] = l->innerGeom[i];
    q->outerGeom[i] = l->outerGeom[i];
  }
  q->nSitesOuter = l->nSitesOuter;
  q->nSitesInner = l->nSitesInner;
  q->nEvenOuter = l->nEvenOuter;
  q->innerCb = l->innerCb;
  q->innerCbDir = l->innerCbDir;
}

void qllFatal(const char *what)
{
  printf("qll_amd: %s failed: %s\n", what, qamd_last_error());
  fflush(stdout);
  exit(-1);
}

/* ------------------------------------------------------------ registry */
typedef struct Reg {
  const void *key;
  void *obj;
  struct Reg *next;
} Reg;
static Reg *g_reg = NULL;

void qllRegPut(const void *key, void *obj)
{
  Reg *r = malloc(sizeof(Reg));
  r->key = key;
  r->obj = obj;
  r->next = g_reg;
  g_reg = r;
}

void *qllRegGet(const void *key)
{
  for (Reg *r = g_reg; r; r = r->next)
    if (r->key == key) return r->obj;
  return NULL;
}

void *qllRegTake(const void *key)
{
  Reg **pp = &g_reg;
  while (*pp) {
    if ((*pp)->key == key) {
      Reg *r = *pp;
      void *o = r->obj;
      *pp = r->next;
      free(r);
      return o;
    }
    pp = &(*pp)->next;
  }
  return NULL;
}

/* one device lattice per distinct geometry, kept for the process */
qamd_lattice *qllLattice(const Layout *l)
{
  static struct { int d[4]; qamd_lattice *lat; } tab[64];
  static int n = 0;
  for (int i = 0; i < n; i++)
    if (!memcmp(tab[i].d, l->physGeom, 4 * sizeof(int))) return tab[i].lat;
  if (l->nDim != 4) {
    printf("qll_amd: only 4-dimensional lattices are supported (nDim = %d)\n", l->nDim);
    exit(-1);
  }
  qamd_lattice *lat = qamd_lattice_create(l->physGeom);
  if (!lat) qllFatal("lattice_create");
  if (n < 64) {
    memcpy(tab[n].d, l->physGeom, 4 * sizeof(int));
    tab[n].lat = lat;
    n++;
  }
  return lat;
}
