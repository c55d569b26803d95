/*
 * layout.h -- QLL site layout and shift types (drop-in for include/layout.h
 * of QLL).  The structs keep QLL's member order and types so callers that
 * read them (l->nSites, sub.beginOuter, ...) work unchanged.  QMP handle
 * members are opaque pointers here (libqll_amd runs on one GPU and sends no
 * messages).
 */
#ifndef QLL_LAYOUT_H
#define QLL_LAYOUT_H
#include <stddef.h>

#ifndef QLL_HAVE_QMP
typedef void *QMP_msgmem_t;
typedef void *QMP_msghandle_t;
#endif

typedef struct llist {
  void *value;
  struct llist *next;
} llist;

/* include/layout.h:10-31 */
typedef struct {
  int nDim;
  int *physGeom;
  int *rankGeom;
  int *innerGeom;
  int *outerGeom;
  int *localGeom;
  int physVol;
  int nEven;
  int nOdd;
  int nSites;
  int nEvenOuter;
  int nOddOuter;
  int nSitesOuter;
  int nSitesInner;
  int innerCb;
  int innerCbDir;
  llist *shifts;
  int nranks;
  int myrank;
} Layout;

typedef struct {
  int rank;
  int index;
} LayoutIndex;

typedef struct {
  int begin;
  int end;
  int beginOuter;
  int endOuter;
} Subset;

/* lib/layout.c:5-217 */
void layoutSetup(Layout *l);
void layoutIndex(Layout *l, LayoutIndex *li, int coords[]);
void layoutCoord(Layout *l, int coords[], LayoutIndex *li);
void layoutShift(Layout *l, LayoutIndex *li, LayoutIndex *li2, int disp[]);
void layoutSubset(Subset *s, Layout *l, char *sub);

typedef struct {
  int myRank;
  int nIndices;
  int *srcRanks;
  int *srcIndices;
  int nRecvDests;
  int nSendIndices;
  int *sendSrcIndices;
  int *sendDestRanks;
  int *sendDestIndices;
} GatherDescription;

typedef struct {
  GatherDescription *gd;
  int myRank;
  int nIndices;
  int *srcIndices;
  int nRecvRanks;
  int *recvRanks;
  int *recvRankSizes;
  int *recvRankOffsets;
  int recvSize;
  int nRecvDests;
  int *recvDestIndices;
  int *recvBufIndices;
  int nSendRanks;
  int *sendRanks;
  int *sendRankSizes;
  int *sendRankOffsets;
  int sendSize;
  int nSendIndices;
  int *sendIndices;
} GatherIndices;

/* shift descriptor (include/layout.h:70-112).  libqll_amd fills disp (the
 * displacement: a shift copies dst(x) = src(x - disp), lib/shifts.c:328) and
 * leaves the message-passing members empty. */
typedef struct {
  GatherIndices *gi;
  int *disp;
  int *sidx;
  int *pidx;
  int nRecvRanks;
  int *recvRanks;
  int *recvRankSizes;
  int *recvRankSizes1;
  int *recvRankOffsets;
  int *recvRankOffsets1;
  int nRecvSites;
  int nRecvSites1;
  int nRecvDests;
  int *recvDests;
  int *recvLocalSrcs;
  int *recvRemoteSrcs;
  int nSendRanks;
  int *sendRanks;
  int *sendRankSizes;
  int *sendRankSizes1;
  int *sendRankOffsets;
  int *sendRankOffsets1;
  int nSendSites;
  int nSendSites1;
  int *sendSites;
  int vv;
  int offr, lenr, nthreads;
  int perm;
  int pack;
  int blend;
  QMP_msgmem_t sqmpmem;
  QMP_msghandle_t smsg;
  QMP_msgmem_t rqmpmem;
  QMP_msghandle_t rmsg;
  QMP_msghandle_t pairmsg;
} ShiftIndices;

typedef struct {
  QMP_msgmem_t sqmpmem;
  QMP_msghandle_t smsg;
  QMP_msgmem_t rqmpmem;
  QMP_msghandle_t rmsg;
  QMP_msghandle_t pairmsg;
  void *sbuf;
  void *rbuf;
  int sbufSize;
  int rbufSize;
  int first;
  int *offr;
  int *lenr;
} ShiftBuf;

#endif
