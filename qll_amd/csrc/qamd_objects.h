// Definitions of the opaque handles of qll_amd.h.
#pragma once
#include "qamd_internal.h"

struct qamd_lattice {
  qamd::Geom g;
};

struct qamd_field {
  qamd::Field f;
  qamd_lattice *lat;
};

// FP64 gauge field, layout [mu][parity][k 0..8][vh_pad] of double2
struct qamd_gauge {
  qamd_lattice *lat;
  double2 *d;
};

// Dslash link set for both output parities; data[p] holds the terms used by
// output sites of parity p in the blocked layout of qamd_internal.h.
struct qamd_links {
  qamd_lattice *lat;
  int kind;     // 0 staggered (doublestored), 1 Wilson, 2 staggered forward-only
  int prec;
  int nterms;
  void *data[2];
  float scale[16];   // 16-bit only: per-slot scale (stored = link / scale)
  // links whose rebuildable terms (the long terms of a 16-term staggered set,
  // every Wilson term, the long slots 4-7 of a forward-only set) are all a
  // scale times a unitary matrix, checked in FP64: the Dslash loads their
  // first two rows only and rebuilds the third; lsign[p][cb] bit k =
  // determinant sign of rebuildable term k
  int recon;
  uint8_t *lsign[2];
  // content version: a process-wide counter value, new after every import,
  // doublestore or conversion into these links
  uint64_t ver;
  // the last sloppy-operator error measured with these links as the sloppy
  // set (qamd_stag_sloppy_error): against links of version err_ref_ver, at
  // mass err_mass, with these links at version err_ver
  mutable uint64_t err_ref_ver, err_ver;
  mutable double err_mass, err_eps;
};

inline bool is_stag(const qamd_links *l) { return l->kind == 0 || l->kind == 2; }
// staggered with long (Naik) terms: 16 doublestored or 8 forward-only terms
inline bool stag_naik(const qamd_links *l) {
  return (l->kind == 0 && l->nterms == 16) || (l->kind == 2 && l->nterms == 8);
}
// two staggered link sets of the same operator form (both with or both
// without Naik terms), stored doublestored or forward-only
inline bool stag_pair(const qamd_links *a, const qamd_links *b) {
  return is_stag(a) && is_stag(b) && stag_naik(a) == stag_naik(b);
}

namespace qamd {
inline double2 *gauge_ptr(const qamd_gauge *u, int mu, int par) {
  return u->d + ((int64_t)(mu * 2 + par) * 9) * u->lat->g.vh_pad;
}
size_t prec_bytes(int prec);   // bytes of one stored complex number
}  // namespace qamd
