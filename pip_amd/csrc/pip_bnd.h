// pip_bnd.h -- ghost-cell index maps of PIPbnd for one rank
// (boundary_control, Boundary_rot.f90:323-402): periodic (type 1), mirror
// (type 2, flag_bnd 2-7 and 11, with per-variable sign maps) and free /
// outflow (type 10), margin 2.  Pure index arithmetic, shared by the HIP
// kernels and the host checks.
#pragma once

#include "pip_phys.h"

namespace PIP_NS {

// Largest padded extent (ix or jx) the packed image list can address.
constexpr int kMaxPad = 32767;

struct Bnd {
  int type[4];      // x-, x+, y-, y+ (1, 2 or 10)
  unsigned sym[4];  // sign bits over NVT variables (U_m planes then U_h planes)
};

// PIPbnd boundary type of a flag_bnd value: 1 periodic, 2 mirror (flags
// 2-7 and 11), 10 free.
PIP_HD int bnd_kind(int f) {
  if ((f >= 3 && f <= 7) || f == 11) return 2;
  return f;
}

// set_sym_mhd / set_sym_hd (Boundary_rot.f90:239-320) as sign bit masks over
// the nv + 5 variables (U_m planes then U_h planes), per side x-, x+, y-, y+.
PIP_HD Bnd make_bnd(const int* flag_bnd, int ndim, int nv, int flag_divb) {
  Bnd b;
  for (int s = 0; s < 4; s++) {
    b.type[s] = 0;
    b.sym[s] = 0u;
  }
  for (int i = 1; i <= 2 * ndim; i++) {
    const int dir = (i + 1) / 2, f = flag_bnd[i - 1];
    int sm[9], sh[5];
    for (int n = 0; n < 9; n++) sm[n] = 1;
    for (int n = 0; n < 5; n++) sh[n] = 1;
    if (f == 2) { sm[dir] = -1; sm[5 + dir - 1] = -1; }
    else if (f == 3) { sm[dir] = -1; sm[6 + dir % 3 - 1] = -1; sm[2 + (dir + 1) % 3 - 1] = -1; }
    else if (f == 4) { sm[dir] = -1; sm[6 + (dir + 1) % 3 - 1] = -1; sm[2 + dir % 3 - 1] = -1; }
    else if (f == 5) { sm[dir] = -1; sm[5 + dir - 1] = -1; sm[2 + (dir + 1) % 3 - 1] = -1; sm[6 + (dir + 1) % 3 - 1] = -1; }
    else if (f == 6) { sm[dir] = -1; sm[6 + dir % 3 - 1] = -1; sm[6 + (dir + 1) % 3 - 1] = -1; }
    else if (f == 7) { sm[1] = -1; sm[2] = -1; sm[3] = -1; }
    if (flag_divb == 1) {
      if (f == 2 || f == 5) sm[8] = 1;
      else if (f == 3 || f == 4 || f == 6 || f == 7) sm[8] = -1;
    }
    if (f == 2) sh[dir] = -1;
    else if (f == 3) { sh[dir] = -1; sh[2 + (dir + 1) % 3 - 1] = -1; }
    else if (f == 4) { sh[dir] = -1; sh[2 + dir % 3 - 1] = -1; }
    else if (f == 5) { sh[dir] = -1; sh[2 + (dir + 1) % 3 - 1] = -1; }
    else if (f == 6) sh[dir] = -1;
    unsigned m = 0u;
    for (int n = 0; n < nv; n++)
      if (sm[n] < 0) m |= 1u << n;
    for (int n = 0; n < 5; n++)
      if (sh[n] < 0) m |= 1u << (nv + n);
    b.type[i - 1] = bnd_kind(f);
    b.sym[i - 1] = m;
  }
  return b;
}

// Source of ghost index g (0,1 or npad-2,npad-1) along one direction.
PIP_HD int src_index(int g, int npad, int tlo, int thi, bool& flip) {
  flip = false;
  if (g < 2) {
    if (tlo == 1) return npad - 4 + g;
    if (tlo == 2) { flip = true; return 3 - g; }
    return 2;
  }
  if (g >= npad - 2) {
    if (thi == 1) return g - (npad - 4);
    if (thi == 2) { flip = true; return 2 * npad - 5 - g; }
    return npad - 3;
  }
  return g;
}

// Ghost images of interior index i along one direction (the inverse of
// src_index).  At most four images (two per side); image k is packed in
// bits 16k..16k+15 as (dst + 1) | (mirror << 15), 0 = no image, so the list
// stays in one register pair.  Needs npad <= kMaxPad.
PIP_HD unsigned long long ghost_images(int i, int npad, int tlo, int thi) {
  unsigned long long out = 0ull;
  int k = 0;
  auto put = [&](int dst, bool flip) {
    out |= ((unsigned long long)(dst + 1) | (flip ? 0x8000ull : 0ull)) << (16 * k);
    k++;
  };
  if (tlo == 1) {
    if (i == npad - 4) put(0, false);
    if (i == npad - 3) put(1, false);
  } else if (tlo == 2) {
    if (i == 3) put(0, true);
    if (i == 2) put(1, true);
  } else if (tlo == 10) {
    if (i == 2) { put(0, false); put(1, false); }
  }
  if (thi == 1) {
    if (i == 2) put(npad - 2, false);
    if (i == 3) put(npad - 1, false);
  } else if (thi == 2) {
    if (i == npad - 3) put(npad - 2, true);
    if (i == npad - 4) put(npad - 1, true);
  } else if (thi == 10) {
    if (i == npad - 3) { put(npad - 2, false); put(npad - 1, false); }
  }
  return out;
}

// Image k of a packed list: false when there is none.
PIP_HD bool ghost_image(unsigned long long g, int k, int& dst, bool& flip) {
  const unsigned e = (unsigned)((g >> (16 * k)) & 0xffffull);
  if (e == 0u) return false;
  dst = (int)(e & 0x7fffu) - 1;
  flip = (e & 0x8000u) != 0u;
  return true;
}

// Visit every ghost image of interior cell (ci, cj) of a 2-D grid (x images
// in the own row, y images in the own column, and the corner images):
// fn(xi, yj, sign_mask).  In 1-D pass cj = 0 and jx = 1 (no y images).
template <typename Fn>
PIP_HD void for_each_image(int ci, int cj, int ix, int jx, const Bnd& b, Fn fn) {
  const unsigned long long gx = ghost_images(ci, ix, b.type[0], b.type[1]);
  const unsigned long long gy = (jx > 1) ? ghost_images(cj, jx, b.type[2], b.type[3]) : 0ull;
  if (gx == 0ull && gy == 0ull) return;
  for (int p = -1; p < 4; p++) {
    int xi = ci;
    unsigned mx = 0u;
    if (p >= 0) {
      bool f;
      if (!ghost_image(gx, p, xi, f)) break;
      mx = f ? b.sym[(xi < 2) ? 0 : 1] : 0u;
    }
    for (int q = -1; q < 4; q++) {
      int yj = cj;
      unsigned my = 0u;
      if (q < 0) {
        if (p < 0) continue;
      } else {
        bool f;
        if (!ghost_image(gy, q, yj, f)) break;
        my = f ? b.sym[(yj < 2) ? 2 : 3] : 0u;
      }
      fn(xi, yj, mx ^ my);
    }
  }
}

}  // namespace PIP_NS
