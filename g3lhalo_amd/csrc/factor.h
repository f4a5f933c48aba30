// factor.h -- compile-time structure of the Genz-Malik point set, used to
// evaluate each distinct sub-expression of an integrand once per region.
//
// Along every dimension the points of one region take only seven
// coordinates: the centre, c -/+ lambda2 h, c -/+ lambda4 h, c -/+ lambda5 h
// (src/hcubature.cpp:343-414).  A quantity that depends on a subset S of the
// dimensions (u_NFW(k1, m, z) depends on l1, m, z) therefore takes at most
// as many distinct values as there are distinct projections of the points
// onto S -- 33 of the 93 points of a 5-D region for a 3-dimensional S.
// Proj<D, MASK> lists those projection classes: cls[q] is the class of
// point q, rep[c] the first point of class c (its representative).  Every
// point of a class has bit-identical coordinates on S, so a value computed
// at the representative is the value the point itself would compute.
#pragma once
#include <stdint.h>

#include "layout.h"

namespace g3l {

// Coordinate code of point q along dimension d, in the point order of
// rule_point: 0 centre, 1 c - l2 h, 2 c + l2 h, 3 c - l4 h, 4 c + l4 h,
// 5 c - l5 h, 6 c + l5 h.
constexpr int gm_code(int D, int q, int d) {
  if (q == 0) return 0;
  int t = q - 1;
  if (t < 4 * D) {
    int k = t >> 2, s = t & 3;
    return k != d ? 0 : 1 + s;
  }
  t -= 4 * D;
  if (t < 2 * D * (D - 1)) {
    int p = t >> 2, s = t & 3;
    int i = 0, rem = p;
    while (rem >= D - 1 - i) {
      rem -= D - 1 - i;
      ++i;
    }
    int j = i + 1 + rem;
    if (d == i) return (s == 1 || s == 2) ? 4 : 3;
    if (d == j) return s >= 2 ? 4 : 3;
    return 0;
  }
  t -= 2 * D * (D - 1);
  unsigned signs = (unsigned)t ^ ((unsigned)t >> 1);
  return ((signs >> d) & 1u) ? 5 : 6;
}

constexpr unsigned bit(int d) { return 1u << d; }

template <int D, unsigned MASK>
struct Proj {
  static constexpr int NP = rule_points(D);
  int n = 0;
  uint8_t cls[NP] = {};
  uint8_t rep[NP] = {};
};

// Key of point q on the dimensions of MASK: three bits per dimension.
constexpr unsigned gm_key(int D, int q, unsigned mask) {
  unsigned k = 0;
  for (int d = 0; d < D; ++d)
    if ((mask >> d) & 1u) k |= (unsigned)gm_code(D, q, d) << (3 * d);
  return k;
}

template <int D, unsigned MASK>
constexpr Proj<D, MASK> make_proj() {
  Proj<D, MASK> p{};
  unsigned key[Proj<D, MASK>::NP] = {};
  unsigned ckey[Proj<D, MASK>::NP] = {};
  for (int q = 0; q < Proj<D, MASK>::NP; ++q) key[q] = gm_key(D, q, MASK);
  for (int q = 0; q < Proj<D, MASK>::NP; ++q) {
    int found = -1;
    for (int c = 0; c < p.n; ++c)
      if (ckey[c] == key[q]) {
        found = c;
        break;
      }
    if (found < 0) {
      p.rep[p.n] = (uint8_t)q;
      ckey[p.n] = key[q];
      found = p.n++;
    }
    p.cls[q] = (uint8_t)found;
  }
  return p;
}

// Device copy of each table (constant memory, indexed per thread).
template <int D, unsigned MASK>
struct ProjTab {
  static constexpr Proj<D, MASK> host = make_proj<D, MASK>();
  static constexpr int n = host.n;
};
template <int D, unsigned MASK>
__constant__ constexpr Proj<D, MASK> kProj = make_proj<D, MASK>();

}  // namespace g3l
