// Matsubara-axis linear-tetrahedron polarization weights on gfx950 (FP64).
//
// Computes PyGW's kcw (bz_calcqdepw_par2, for_tetrahedra.f90:390-464):
//   kcw[n,m,k,w] = sum over tetrahedra t with corner k of
//                  6/N_tet * w_corner(t,k; e_n(t), e_m(t linked by q), w)
// with the corner weights of convw1t_par (for_tetrahedra3.f90:3-136).
//
// Design (deterministic, no floating-point atomics):
//   * pass 0: the atan / log of Delta = e_f(k-q) - e_e(k) for every
//     (k-point, frequency, pair) of the chunk, once -- a whole tetrahedron's
//     vertex energies are these per-k values, so each is shared by the 24
//     tetrahedra around k and by their four corners;
//   * pass 1, one lane per (tetrahedron, band pair), a wavefront holding 64
//     pairs of one tetrahedron: the four corner weights over all Matsubara
//     frequencies (the reference's w1t(4, nom) of one convw1t_par call).  For
//     the whole tetrahedron (no E_F plane inside) each corner weight is one
//     vertex's closed form.  sorteq classes 16, 8 and 10 (nearly equal
//     corner Delta's: most tetrahedra on a fine mesh) share their averaged
//     vertex energies, so one frequency costs at most two atan / log for all
//     four corners; classes 4 and 6 are compacted into a list evaluated per
//     (pair, corner) on the tabulated atan / log (used only where the vertex
//     energy equals the table's bit for bit), and polyhedra cut by E_F into
//     another list evaluated by four cooperating lanes per pair on the
//     general path, so neither stalls the wavefronts of whole tetrahedra;
//   * pass 2, one lane per (k, frequency, band pair): the 24 (tetrahedron, corner)
//     incidences of k summed in increasing tetrahedron order -- the order in
//     which the reference scatters into kcw(:,:,k,:), so the FP64 sums
//     associate identically; writes coalesced across the wavefront (pair
//     index fastest in both output layouts);
//   * the pass-1 buffer holds ntet*4*nom doubles per pair; pairs are taken in
//     chunks sized to the workspace.
// Arithmetic follows the reference expression by expression (no FMA
// contraction in this file); atan/log come from the device math library, so
// results agree with the CPU reference to the conditioning of the closed
// forms, not bit for bit.
#include "common.h"

#pragma clang fp contract(off)

namespace tw {

constexpr int OMC = 8;           // Matsubara points per register block
constexpr int MAXN = 20;         // candidate polyhedron nodes

// compatibility switches (bit set = reproduce the reference)
constexpr int COMPAT_F7B = 1;    // case-4 vertex rotation (for_tetrahedra3.f90:482-486)
constexpr int COMPAT_N78 = 2;    // wtemp(4,2,nom) aliasing for nnod = 7, 8 (for_tetrahedra3.f90:33,105-118)

// status bits written to the device status word
constexpr int ST_GEOM = 1;       // a geometry/classification error (reference `stop`)
constexpr int ST_F7A = 2;        // nnod=5 bipyramid (reference undefined)

__device__ __forceinline__ double sq(double x) { return x * x; }
__device__ __forceinline__ double cu(double x) { return (x * x) * x; }
__device__ __forceinline__ double qu(double x) { return ((x * x) * x) * x; }

// ---------------------------------------------------------------------------
// node geometry: surfnodes / unrepnodes / relnodes / sortnodes
// (for_tetrahedra2.f90:1803-2773), restated for one lane
// ---------------------------------------------------------------------------
struct Nodes {
  double x[MAXN][3];
  int ty[MAXN];
  int n;
};

__device__ void surf_nodes(Nodes& g, const double* e, const double* f, double ef) {
  const double zt = 1.0e-12;
  const int code[20] = {7, 11, 13, 14, 19, 21, 22, 25, 26, 28, 35, 37, 38, 41, 42, 44, 49, 50, 52, 56};
  // corners (0,0,0),(0,0,1),(0,1,0),(1,0,0)
  for (int i = 0; i < 4; i++) { g.x[i][0] = g.x[i][1] = g.x[i][2] = 0.0; g.ty[i] = code[i]; }
  g.x[1][2] = 1.0; g.x[2][1] = 1.0; g.x[3][0] = 1.0;
  int n = 4, iint = 4;
  auto put = [&](double a, double b, double c, int t) {
    g.x[n][0] = a; g.x[n][1] = b; g.x[n][2] = c; g.ty[n] = t; n++;
  };
  // plane e=ef with the three edges from corner 1 (nodes 5..7)
  for (int j = 2; j <= 4; j++) {
    int i = 6 - j;  // 4,3,2
    iint++;
    double den = e[i - 1] - e[0];
    if (fabs(den) > zt) {
      double t = (ef - e[0]) / den;
      if (t > 0.0 && t <= 1.0) {
        double v[3] = {0.0, 0.0, 0.0};
        v[i - 2] = t;
        put(v[0], v[1], v[2], code[iint - 1]);
      }
    }
  }
  // plane e=ef with the edges of face 4 (nodes 8..10)
  for (int k = 1; k <= 3; k++) {
    int i = k % 3 + 1, j = i % 3 + 1;
    iint++;
    double den = e[i] - e[j];
    if (fabs(den) > zt) {
      double v[3] = {0.0, 0.0, 0.0};
      v[i - 1] = (ef - e[j]) / den;
      v[j - 1] = (e[i] - ef) / den;
      if (v[i - 1] >= 0.0 && v[j - 1] >= 0.0) put(v[0], v[1], v[2], code[iint - 1]);
    }
  }
  // plane f=ef with the edges from corner 1 (nodes 11..13)
  for (int j = 2; j <= 4; j++) {
    int i = 6 - j;
    iint++;
    double den = f[i - 1] - f[0];
    if (fabs(den) > zt) {
      double t = (ef - f[0]) / den;
      if (t >= 0.0 && t <= 1.0) {
        double v[3] = {0.0, 0.0, 0.0};
        v[i - 2] = t;
        put(v[0], v[1], v[2], code[iint - 1]);
      }
    }
  }
  // plane f=ef with the edges of face 4 (nodes 14..16)
  for (int k = 1; k <= 3; k++) {
    int i = k % 3 + 1, j = i % 3 + 1;
    iint++;
    double den = f[i] - f[j];
    if (fabs(den) > zt) {
      double v[3] = {0.0, 0.0, 0.0};
      v[i - 1] = (ef - f[j]) / den;
      v[j - 1] = (f[i] - ef) / den;
      if (v[i - 1] >= 0.0 && v[j - 1] >= 0.0) put(v[0], v[1], v[2], code[iint - 1]);
    }
  }
  // edif(a,b)*fdif(c,b) - edif(c,b)*fdif(a,b), 1-based a,b,c
  auto efdif = [&](int a, int b, int c) {
    return (e[a - 1] - e[b - 1]) * (f[c - 1] - f[b - 1]) - (e[c - 1] - e[b - 1]) * (f[a - 1] - f[b - 1]);
  };
  auto efef = [&](int a, int b) {
    return (ef - e[a - 1]) * (f[b - 1] - f[a - 1]) - (ef - f[a - 1]) * (e[b - 1] - e[a - 1]);
  };
  // both planes with faces 1..3 (nodes 17..19)
  for (int k = 1; k <= 3; k++) {
    int i = k % 3 + 1, j = i % 3 + 1;
    iint++;
    double d1 = efdif(i + 1, 1, j + 1), d2 = efdif(j + 1, 1, i + 1);
    if (fabs(d1) > zt && fabs(d2) > zt) {
      double v[3] = {0.0, 0.0, 0.0};
      v[i - 1] = efef(1, j + 1) / d1;
      v[j - 1] = efef(1, i + 1) / d2;
      double s = v[i - 1] + v[j - 1];
      if (v[i - 1] > 0.0 && v[j - 1] > 0.0 && s <= 1.0) put(v[0], v[1], v[2], code[iint - 1]);
    }
  }
  // both planes with face 4 (node 20)
  {
    double d1 = efdif(2, 4, 3), d2 = efdif(3, 4, 2), d3 = efdif(4, 2, 3);
    iint++;
    if (fabs(d1) > zt && fabs(d2) > zt && fabs(d3) > zt) {
      double a = efef(4, 3) / d1, b = efef(4, 2) / d2, c = efef(2, 3) / d3;
      if (a > 0.0 && b > 0.0 && c > 0.0) put(a, b, c, code[iint - 1]);
    }
  }
  g.n = n;
}

__device__ void unrep_nodes(Nodes& g) {
  const double zt = 1.0e-8;
  int n = g.n;
  for (int i = 0; i < n - 1; i++) {
    double a0 = g.x[i][0], a1 = g.x[i][1], a2 = g.x[i][2];
    int j = i + 1;
    while (j < n) {
      double s = 0.0;
      s = s + fabs(a0 - g.x[j][0]);
      s = s + fabs(a1 - g.x[j][1]);
      s = s + fabs(a2 - g.x[j][2]);
      if (s < zt) {
        g.ty[i] |= (g.ty[j] & 63);
        for (int k = j + 1; k < n; k++) {
          g.x[k - 1][0] = g.x[k][0]; g.x[k - 1][1] = g.x[k][1]; g.x[k - 1][2] = g.x[k][2];
          g.ty[k - 1] = g.ty[k];
        }
        n--;
      } else {
        j++;
      }
    }
  }
  g.n = n;
}

__device__ __forceinline__ double tlin(const double* x, const double* v) {
  double fx = v[0];
  fx = fx + (v[1] - v[0]) * x[0];
  fx = fx + (v[2] - v[0]) * x[1];
  fx = fx + (v[3] - v[0]) * x[2];
  return fx;
}

__device__ bool rel_nodes(Nodes& g, const double* e, const double* f, double ef) {
  const double zt = 1.0e-12;
  bool ok = true;
  int i = 0, n = g.n;
  while (i < n) {
    double e1 = tlin(g.x[i], e) - ef;
    double e2 = tlin(g.x[i], f) - ef;
    int t = g.ty[i];
    bool keep = false;
    bool occ = (e1 < 0.0) || (fabs(e1) <= zt);
    bool emp = (e2 > 0.0) || (fabs(e2) <= zt);
    if (t >= 7 && t <= 15) keep = occ && emp;
    else if (t >= 16 && t <= 31) keep = emp;
    else if (t >= 32 && t <= 47) keep = occ;
    else if (t >= 48 && t <= 63) keep = true;
    else ok = false;
    if (keep) {
      i++;
    } else {
      for (int j = i + 1; j < n; j++) {
        g.x[j - 1][0] = g.x[j][0]; g.x[j - 1][1] = g.x[j][1]; g.x[j - 1][2] = g.x[j][2];
        g.ty[j - 1] = g.ty[j];
      }
      n--;
    }
  }
  g.n = n;
  return ok;
}

__device__ __forceinline__ bool bt(int x, int b) { return (x >> b) & 1; }

// Decomposition positions: pos[inode][sub] (1-based slot, 0 = absent).
// Returns false where the reference stops.  Same branch structure as
// sortnodes (for_tetrahedra2.f90:2306-2773); the nnod=5 six-triangle case
// uses the intended (not the reference's undefined) counters and sets ST_F7A.
__device__ bool sort_nodes(const Nodes& g, int8_t pos[MAXN][2], int* status) {
  const int n = g.n;
  for (int i = 0; i < MAXN; i++) pos[i][0] = pos[i][1] = 0;
  if (n < 5 || n > 8) return true;
  int spl[6];
  for (int b = 0; b < 6; b++) {
    int c = 0;
    for (int i = 0; i < n; i++) c += bt(g.ty[i], b);
    spl[b] = c;
  }
  int tri[6] = {0, 0, 0, 0, 0, 0}, sqr[6] = {0, 0, 0, 0, 0, 0}, pen[2] = {0, 0};
  int it = 0, is = 0, ip = 0;
  if (n == 5) {
    int sb = 0;
    for (int b = 0; b < 6; b++) {
      if (spl[b] == 3) tri[it++] = b;
      else if (spl[b] == 4) { sqr[is++] = b; sb = b; }
    }
    if (is != 0) {
      int is1 = 1, sn = 0;
      for (int i = 0; i < 5; i++)
        if (bt(g.ty[i], tri[0])) {
          if (!bt(g.ty[i], sb)) pos[i][0] = 2;
          else { pos[i][0] = is1; is1 += 2; sn = i; }
        }
      for (int t = 1; t < 4; t++) if (bt(g.ty[sn], tri[t])) sb = tri[t];
      for (int i = 0; i < 5; i++) if (pos[i][0] == 0) pos[i][0] = bt(g.ty[i], sb) ? 5 : 4;
    } else {
      int it1 = 1, it2 = 1;
      *status |= ST_F7A;
      for (int i = 0; i < 5; i++) {
        int nb = 0;
        for (int b = 0; b < 6; b++) nb += bt(g.ty[i], b);
        if (nb == 3) { pos[i][0] = 1 + (it1 - 1) * 4; it1++; }
        else { pos[i][0] = 1 + it2; it2++; }
      }
      if (it1 != 3 || it2 != 4) return false;
    }
  } else if (n == 6) {
    for (int b = 0; b < 6; b++) {
      if (spl[b] == 3) tri[it++] = b;
      else if (spl[b] == 4) sqr[is++] = b;
    }
    if (is == 2) {
      if (it != 4) return false;
      int cnt = 1, sn = -1, snb = -1;
      for (int i = 0; i < 6; i++)
        if (bt(g.ty[i], sqr[0]) && bt(g.ty[i], sqr[1])) {
          pos[i][0] = 2 + (cnt - 1) * 3;
          if (cnt == 1) sn = i;
          if (cnt == 2) snb = i;
          cnt++;
        }
      if (sn < 0 || snb < 0) return false;
      int sa = -1, sbb = -1;
      for (int t = 0; t < 4; t++) {
        if (bt(g.ty[sn], tri[t])) sa = t;
        if (bt(g.ty[snb], tri[t])) sbb = t;
      }
      if (sa < 0 || sbb < 0) return false;
      for (int i = 0; i < 6; i++) {
        if (pos[i][0] != 0) continue;
        int st = 0;
        for (int t = 0; t < 4; t++) st += bt(g.ty[i], tri[t]);
        if (st == 3 && bt(g.ty[i], tri[sa])) pos[i][0] = 4;
        if (st == 2 && bt(g.ty[i], tri[sa])) pos[i][0] = 1;
        if (st == 3 && bt(g.ty[i], tri[sbb])) pos[i][0] = 3;
        if (st == 2 && bt(g.ty[i], tri[sbb])) pos[i][0] = 6;
      }
    } else if (is == 3) {
      for (int i = 0; i < 6; i++)
        for (int t = 0; t < 2; t++)
          if (bt(g.ty[i], tri[t])) {
            if (bt(g.ty[i], sqr[0]) && bt(g.ty[i], sqr[1])) pos[i][0] = 1 + 3 * t;
            else if (bt(g.ty[i], sqr[0])) pos[i][0] = 2 + 3 * t;
            else if (bt(g.ty[i], sqr[1])) pos[i][0] = 3 + 3 * t;
            else return false;
          }
    } else {
      return false;
    }
  } else if (n == 7) {
    for (int b = 0; b < 6; b++) {
      if (spl[b] == 3) tri[it++] = b;
      else if (spl[b] == 4) sqr[is++] = b;
      else if (spl[b] == 5) { if (ip < 2) pen[ip] = b; ip++; }
    }
    if (it == 2) {
      if (is != 4) return false;
      int sn = -1;
      for (int i = 0; i < 7; i++)
        if (bt(g.ty[i], tri[0]) && bt(g.ty[i], tri[1])) { pos[i][0] = pos[i][1] = 2; sn = i; }
      if (sn < 0) return false;
      int is1 = 0, sa = -1, sbb = -1;
      for (int b = 0; b < 6; b++) {
        if (bt(g.ty[sn], b) && spl[b] == 4 && is1 == 0) { sa = b; is1++; }
        else if (bt(g.ty[sn], b) && spl[b] == 4) sbb = b;
      }
      if (sa < 0 || sbb < 0) return false;
      for (int i = 0; i < 7; i++) {
        if (i == sn) continue;
        int t = g.ty[i];
        if (bt(t, tri[0]) && bt(t, sa)) pos[i][0] = 1;
        else if (bt(t, tri[0]) && bt(t, sbb)) pos[i][0] = 3;
        else if (bt(t, tri[1]) && bt(t, sa)) pos[i][1] = 1;
        else if (bt(t, tri[1]) && bt(t, sbb)) pos[i][1] = 3;
        else if (bt(t, sa)) { pos[i][0] = pos[i][1] = 4; }
        else { pos[i][0] = pos[i][1] = 5; }
      }
    } else if (it == 3) {
      if (ip < 1 || is < 2) return false;
      int iind = -1;
      for (int i = 0; i < 7; i++) {
        int nb = 0;
        for (int t = 0; t < 3; t++) nb += bt(g.ty[i], tri[t]);
        if (bt(g.ty[i], pen[0])) {
          if (nb == 2) pos[i][0] = 1;
        } else {
          if (nb == 2) { pos[i][0] = 2; pos[i][1] = 1; }
          else { pos[i][1] = 4; iind = i; }
        }
      }
      if (iind < 0) return false;
      int sa = -1;
      for (int t = 0; t < 3; t++) if (bt(g.ty[iind], tri[t])) sa = tri[t];
      if (sa < 0) return false;
      for (int i = 0; i < 7; i++) {
        if (pos[i][0] != 0 || pos[i][1] != 0) continue;
        int t = g.ty[i];
        if (bt(t, sa) && bt(t, sqr[0])) pos[i][1] = 5;
        else if (bt(t, sa) && bt(t, sqr[1])) { pos[i][0] = 5; pos[i][1] = 2; }
        else if (bt(t, sqr[0])) { pos[i][0] = 4; pos[i][1] = 3; }
        else pos[i][0] = 3;
      }
    } else {
      return false;
    }
  } else {  // n == 8
    for (int b = 0; b < 6; b++) {
      if (spl[b] == 3) tri[it++] = b;
      else if (spl[b] == 4) sqr[is++] = b;
      else if (spl[b] == 5) { if (ip < 2) pen[ip] = b; ip++; }
    }
    if (is == 2) {
      if (it < 2 || ip < 2) return false;
      for (int i = 0; i < 8; i++) {
        int t = g.ty[i];
        for (int a = 0; a < 2; a++)
          if (bt(t, tri[a])) {
            if (bt(t, pen[0]) && bt(t, pen[1])) pos[i][0] = 1 + a * 3;
            else if (bt(t, pen[0])) pos[i][0] = 2 + a * 3;
            else if (bt(t, pen[1])) pos[i][0] = 3 + a * 3;
          }
        for (int a = 0; a < 2; a++)
          if (bt(t, pen[a])) {
            if (bt(t, sqr[0]) && bt(t, sqr[1])) pos[i][1] = 1 + a * 3;
            else if (bt(t, sqr[0])) pos[i][1] = 2 + a * 3;
            else if (bt(t, sqr[1])) pos[i][1] = 3 + a * 3;
          }
      }
    } else if (is == 6) {
      int is1 = 0, ib[3] = {0, 0, 0};
      pos[0][0] = 1; pos[0][1] = 0;
      for (int b = 0; b < 6; b++) if (bt(g.ty[0], b)) { if (is1 < 3) ib[is1] = b; is1++; }
      if (is1 != 3) return false;
      for (int i = 1; i < 8; i++) {
        int t = g.ty[i];
        if (bt(t, ib[0])) {
          if (bt(t, ib[1])) { pos[i][0] = pos[i][1] = 2; }
          else if (bt(t, ib[2])) { pos[i][0] = pos[i][1] = 3; }
          else { pos[i][0] = 0; pos[i][1] = 1; }
        } else if (bt(t, ib[1])) {
          if (bt(t, ib[2])) { pos[i][0] = 4; pos[i][1] = 0; }
          else { pos[i][0] = pos[i][1] = 5; }
        } else {
          if (bt(t, ib[2])) { pos[i][0] = pos[i][1] = 6; }
          else { pos[i][0] = 0; pos[i][1] = 4; }
        }
      }
    } else {
      return false;
    }
  }
  // every slot of each sub-polyhedron filled exactly once
  int npos = (n == 5 || n == 7) ? 5 : 6, nsub = n > 6 ? 2 : 1;
  for (int s = 0; s < nsub; s++) {
    int seen = 0;
    for (int i = 0; i < n; i++) {
      int p = pos[i][s];
      if (p < 0 || p > npos) return false;
      if (p) { if (seen & (1 << p)) return false; seen |= 1 << p; }
    }
    if (seen != ((1 << (npos + 1)) - 2)) return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// closed forms per vertex (stweight_imag_par, for_tetrahedra3.f90:439-743)
// ---------------------------------------------------------------------------
// kind: 4 (class 4), 6 (class 6), 2 (classes 8 and 10: two-node form), 16
struct VCoef {
  double t[9], w[9];
  double cc, sgn6;
  double ev[4];
  int kind;
  int na, nb;        // two-node form: slots of ev used
};

#define E(i) ev[(i) - 1]
#define V(i, j) (ev[(i) - 1] - ev[(j) - 1])
// class 4 (four distinct energies): coefficients of the vertex whose
// rotated energies are ev
__device__ __forceinline__ void coef_cls4(const double* ev, double* t, double* w, double& cc) {
  double ev3t4 = E(3) * E(4);
  double ev1_2 = E(1) * E(1);
  double e22 = 3.0 * sq(E(2)), e32 = 3.0 * sq(E(3)), e42 = 3.0 * sq(E(4));
  double vdf24 = V(2, 3) * V(2, 4) * V(3, 4);
  w[0] = 2.0 * V(1, 2) * (V(1, 3) * V(1, 4)) * vdf24;
  t[0] = -ev1_2 * w[0];
  w[1] = vdf24 * 2.0 * ((2.0 * E(1) - E(2)) * (E(3) + E(4)) + E(1) * (2.0 * E(2) - 3.0 * E(1)) - E(3) * E(4));
  t[1] = vdf24 * 6.0 * E(1) * (2.0 * E(2) * ev3t4 + E(1) * (ev1_2 - ev3t4 - E(2) * (E(3) + E(4))));
  w[2] = 2.0 * sq(V(1, 3) * V(1, 4)) * V(3, 4);
  t[2] = -e22 * w[2];
  w[3] = -2.0 * sq(V(1, 2) * V(1, 4)) * V(2, 4);
  t[3] = -e32 * w[3];
  w[4] = 2.0 * sq(V(1, 2) * V(1, 3)) * V(2, 3);
  t[4] = -e42 * w[4];
  t[5] = vdf24 * ev1_2 * (3.0 * E(2) * ev3t4 + E(1) * (E(2) * (E(1) - 2.0 * E(3)) + E(4) * (E(1) - 2.0 * E(2)) + E(3) * (E(1) - 2.0 * E(4))));
  w[5] = -3.0 * vdf24 * (E(2) * E(3) * E(4) + ev1_2 * (2.0 * E(1) - E(2) - E(3) - E(4)));
  double d2 = E(2) * sq(V(1, 3) * V(1, 4)) * V(3, 4) / 3.0;
  double d3 = E(3) * sq(V(1, 2) * V(1, 4)) * V(2, 4) / 3.0;
  double d4 = E(4) * sq(V(1, 2) * V(1, 3)) * V(2, 3) / 3.0;
  t[6] = -e22 * d2;
  w[6] = 9.0 * d2;
  w[7] = -9.0 * d3;
  t[7] = e32 * d3;
  w[8] = 9.0 * d4;
  t[8] = -e42 * d4;
  cc = 6.0 * sq(V(1, 2) * V(1, 3) * V(1, 4)) * V(2, 3) * V(2, 4) * V(3, 4);
}

// sorted-vertex v (0..3) of class cls with sorted Delta d
__device__ void vertex_coef(VCoef& c, int cls, const double* d, int v, int compat) {
  double ev[4];
  if (cls == 4) {
    if (compat & COMPAT_F7B) { for (int j = 0; j < 4; j++) ev[(j + v) & 3] = d[j]; }
    else { for (int j = 0; j < 4; j++) ev[j] = d[(j + v) & 3]; }
    coef_cls4(ev, c.t, c.w, c.cc);
    c.kind = 4;
  } else if (cls == 6) {
    if (v < 2) {
      ev[0] = d[v]; ev[1] = ev[0]; ev[2] = d[2]; ev[3] = d[3];
      double evf = V(3, 1) * V(1, 4) * V(3, 4);
      c.w[0] = (2.0 * E(1) - E(3) - E(4)) * evf * 2.0;
      c.t[0] = E(1) * (5.0 * E(3) * E(4) + E(1) * (E(1) - 3.0 * E(3) - 3.0 * E(4))) * evf;
      c.t[1] = -6.0 * (sq(E(1)) * (E(1) * E(3) + E(1) * E(4) - 3.0 * E(3) * E(4)) + sq(E(3) * E(4))) * V(3, 4);
      c.w[1] = 2.0 * (3.0 * E(1) * (E(1) - E(3) - E(4)) + sq(E(3)) + E(3) * E(4) + sq(E(4))) * V(3, 4);
      c.w[2] = -2.0 * cu(V(1, 4));
      c.t[2] = -3.0 * sq(E(3)) * c.w[2];
      c.w[3] = 2.0 * cu(V(1, 3));
      c.t[3] = -3.0 * sq(E(4)) * c.w[3];
      c.t[4] = -V(3, 4) * E(1) * (3.0 * E(3) * E(4) * (E(1) * (E(1) - E(3) - E(4)) + E(3) * E(4)) + sq(E(1)) * sq(E(3) - E(4)));
      c.w[4] = -3.0 * V(3, 4) * (E(3) * E(4) * (3.0 * E(1) - E(3) - E(4)) - cu(E(1)));
      evf = E(3) * cu(V(1, 4));
      c.w[5] = -3.0 * evf;
      c.t[5] = sq(E(3)) * evf;
      evf = E(4) * cu(V(1, 3));
      c.w[6] = 3.0 * evf;
      c.t[6] = -sq(E(4)) * evf;
      c.cc = 6.0 * cu(V(1, 3)) * cu(V(1, 4)) * V(3, 4);
    } else {
      ev[0] = (d[0] + d[1]) * 0.5;
      ev[1] = ev[0];
      if (v == 2) { ev[2] = d[2]; ev[3] = d[3]; } else { ev[2] = d[3]; ev[3] = d[2]; }
      c.t[0] = (sq(E(3)) * E(4) + sq(E(1)) * V(4, 3) - E(1) * sq(E(3))) * 2.0 * V(1, 3) * V(1, 4) * V(3, 4);
      c.w[0] = ((E(1) + E(3) - 2.0 * E(4))) * 2.0 * V(1, 3) * V(1, 4) * V(3, 4);
      c.t[1] = (3.0 * E(1) * (E(1) * (E(3) + E(1)) - 2.0 * E(3) * E(4))) * 2.0 * sq(V(3, 4));
      c.w[1] = (E(3) + 2.0 * E(4) - 3.0 * E(1)) * 2.0 * sq(V(3, 4));
      c.t[2] = -3.0 * E(3) * (sq(E(3)) + E(1) * (E(3) - 2.0 * E(4))) * 2.0 * sq(V(1, 4));
      c.w[2] = -(E(1) - 3.0 * E(3) + 2.0 * E(4)) * 2.0 * sq(V(1, 4));
      c.w[3] = 2.0 * cu(V(1, 3));
      c.t[3] = (-3.0 * sq(E(4))) * c.w[3];
      c.t[4] = sq(E(1)) * (-3.0 * E(3) * E(4) + E(1) * (2.0 * E(3) + E(4))) * sq(V(3, 4));
      c.w[4] = (-6.0 * sq(E(1)) + 3.0 * E(1) * E(4) + 3.0 * E(3) * E(4)) * sq(V(3, 4));
      c.t[5] = -sq(E(3)) * (2.0 * E(1) * E(3) - 3.0 * E(1) * E(4) + E(3) * E(4)) * sq(V(1, 4));
      c.w[5] = -3.0 * (-2.0 * sq(E(3)) + (E(1) + E(3)) * E(4)) * sq(V(1, 4));
      c.t[6] = -E(4) * cu(V(1, 3)) * sq(E(4));
      c.w[6] = E(4) * cu(V(1, 3)) * 3.0;
      c.cc = 6.0 * cu(V(1, 3)) * sq(V(1, 4)) * sq(V(3, 4));
    }
    c.kind = 6;
  } else if (cls == 8) {
    ev[0] = (d[0] + d[1]) / 2.0; ev[1] = ev[0];
    ev[2] = (d[2] + d[3]) / 2.0; ev[3] = ev[2];
    if (v < 2) {
      c.w[0] = 6.0 * V(3, 1);                                              // evw(1)
      c.t[0] = (sq(E(1)) - 5.0 * E(1) * E(3) - 2.0 * sq(E(3))) * V(3, 1);  // evt(1)
      c.t[1] = -6.0 * (sq(E(3)) + 2.0 * E(1) * E(3));                      // evt(2)
      c.w[2] = 3.0 * (E(1) + 2.0 * E(3));                                  // evw(3)
      c.t[2] = -3.0 * E(1) * sq(E(3));                                     // evt(3)
      c.sgn6 = 6.0;
    } else {
      c.w[0] = 6.0 * V(1, 3);                                              // evw(4)
      c.t[0] = -(2.0 * sq(E(1)) + 5.0 * E(1) * E(3) - sq(E(3))) * V(1, 3); // evt(4)
      c.t[1] = 6.0 * (sq(E(1)) + 2.0 * E(1) * E(3));                       // evt(5)
      c.w[2] = -3.0 * (2.0 * E(1) + E(3));                                 // evw(6)
      c.t[2] = 3.0 * sq(E(1)) * E(3);                                      // evt(6)
      c.sgn6 = -6.0;
    }
    c.cc = 6.0 * qu(V(1, 3));
    c.kind = 2; c.na = 0; c.nb = 2;
  } else if (cls == 10) {
    double a = (d[0] + d[1] + d[2]) / 3.0;
    ev[0] = a; ev[1] = a; ev[2] = a; ev[3] = d[3];
    if (v < 3) {
      c.w[0] = 6.0 * V(1, 4);
      c.t[0] = -V(1, 4) * (2.0 * sq(E(1)) - 7.0 * E(1) * E(4) + 11.0 * sq(E(4)));
      c.t[1] = 18.0 * sq(E(4));
      c.w[2] = -9.0 * E(4);
      c.t[2] = 3.0 * cu(E(4));
      c.sgn6 = -6.0;
      c.cc = 18.0 * qu(V(1, 4));
    } else {
      c.w[0] = 6.0 * V(4, 1);
      c.t[0] = V(4, 1) * (sq(E(1)) - 5.0 * E(1) * E(4) - 2.0 * sq(E(4)));
      c.t[1] = -6.0 * E(4) * (E(4) + 2.0 * E(1));
      c.w[2] = 3.0 * (E(1) + 2.0 * E(4));
      c.t[2] = -3.0 * E(1) * sq(E(4));
      c.sgn6 = 6.0;
      c.cc = 6.0 * qu(V(1, 4));
    }
    c.kind = 2; c.na = 0; c.nb = 3;
  } else {  // 16
    ev[0] = (d[0] + d[1] + d[2] + d[3]) / 4.0;
    ev[1] = ev[2] = ev[3] = ev[0];
    c.kind = 16;
  }
  for (int i = 0; i < 4; i++) c.ev[i] = ev[i];
#undef E
#undef V
}

__device__ __forceinline__ void tan_log(double x, double om, double& t, double& l) {
  if (om > F32C(1.0e-20)) t = atan(x / om);
  else t = copysign(1.0, x) * 2.0 * atan(1.0);
  if ((om * om + x * x) > 1.0e-10) l = log(x * x + om * om);
  else l = 0.0;
}

// class-4 closed form at one frequency from the atan / log of the vertex
// energies (T[i], L[i] of ev[i])
__device__ __forceinline__ double eval_cls4(const double* t, const double* w, double cc, const double* T,
                                            const double* L, double om) {
  const double om2 = om * om;
  double aa = t[0] + om2 * w[0] + om * (T[0] * t[1] + T[1] * t[2] + T[2] * t[3] + T[3] * t[4]
             + om2 * (T[0] * w[1] + T[1] * w[2] + T[2] * w[3] + T[3] * w[4]));
  double bb = L[0] * t[5] + L[1] * t[6] + L[2] * t[7] + L[3] * t[8]
             + om2 * (L[0] * w[5] + L[1] * w[6] + L[2] * w[7] + L[3] * w[8]);
  return (aa + bb) / cc;
}

// check_weight (for_tetrahedra3.f90:694-710): |w| > 1e4 is zeroed at omega ~ 0, else counted
__device__ __forceinline__ double check_weight(double w, double om, int* nwarn) {
  if (fabs(w) > 1.0e4) {
    if (fabs(om) < F32C(1.e-6)) w = 0.0;
    else (*nwarn)++;
  }
  return w;
}

// class-6 closed form from T/L of ev[0], ev[2], ev[3]
__device__ __forceinline__ double eval_cls6(const VCoef& c, double om, double T0, double L0, double T2, double L2,
                                            double T3, double L3) {
  const double om2 = om * om;
  double aa = c.t[0] + om2 * c.w[0] + om * (T0 * c.t[1] + T2 * c.t[2] + T3 * c.t[3]
             + om2 * (T0 * c.w[1] + T2 * c.w[2] + T3 * c.w[3]));
  double bb = L0 * c.t[4] + L2 * c.t[5] + L3 * c.t[6] + om2 * (L0 * c.w[4] + L2 * c.w[5] + L3 * c.w[6]);
  return (aa + bb) / c.cc;
}

// two-node closed form (classes 8, 10) from T/L of ev[na], ev[nb]
__device__ __forceinline__ double eval_kind2(const VCoef& c, double om, double Ta, double La, double Tb, double Lb) {
  const double om2 = om * om;
  double aa = c.t[0] + c.w[0] * om2 + om * (Ta - Tb) * (c.sgn6 * om2 + c.t[1]);
  double bb = (La - Lb) * (c.t[2] + c.w[2] * om2);
  return (aa + bb) / c.cc;
}

// class 16 (all four equal to a)
__device__ __forceinline__ double eval_cls16(double a, double om) { return -a / (12.0 * (om * om + a * a)); }

__device__ __forceinline__ double vertex_eval(const VCoef& c, double om, int* nwarn) {
  double w;
  if (c.kind == 4) {
    double T[4], L[4];
    for (int i = 0; i < 4; i++) tan_log(c.ev[i], om, T[i], L[i]);
    w = eval_cls4(c.t, c.w, c.cc, T, L, om);
  } else if (c.kind == 6) {
    double T0, L0, T2, L2, T3, L3;
    tan_log(c.ev[0], om, T0, L0);
    tan_log(c.ev[2], om, T2, L2);
    tan_log(c.ev[3], om, T3, L3);
    w = eval_cls6(c, om, T0, L0, T2, L2, T3, L3);
  } else if (c.kind == 2) {
    double Ta, La, Tb, Lb;
    tan_log(c.ev[c.na], om, Ta, La);
    tan_log(c.ev[c.nb], om, Tb, Lb);
    w = eval_kind2(c, om, Ta, La, Tb, Lb);
  } else {
    w = eval_cls16(c.ev[0], om);
  }
  return check_weight(w, om, nwarn);
}

// stweight_itaylor_par for original vertex j (for_tetrahedra3.f90:407-437)
__device__ __forceinline__ void taylor_coef(const double* d, int j, double& s1, double& n03) {
  double ev[4];
  for (int i = 0; i < 4; i++) ev[i] = d[(i + j) & 3];
  s1 = -(2.0 * ev[0] + ev[1] + ev[2] + ev[3]);
  double n = 4.0 * cu(ev[0]) + 3.0 * sq(ev[0]) * (ev[1] + ev[2] + ev[3]);
  n = n + 2.0 * ev[0] * (sq(ev[1]) + sq(ev[2]) + sq(ev[3]) + ev[1] * ev[2] + ev[1] * ev[3] + ev[2] * ev[3]);
  n = n + cu(ev[1]) + cu(ev[2]) + cu(ev[3]) + sq(ev[1]) * ev[2] + sq(ev[1]) * ev[3] + sq(ev[2]) * ev[3]
        + ev[1] * sq(ev[2]) + ev[1] * sq(ev[3]) + ev[2] * sq(ev[3]) + ev[1] * ev[2] * ev[3];
  n03 = n;
}

// sorteq (for_tetrahedra2.f90:1688-1798): in-place sort of v, ind[sorted] = original
__device__ bool sort_eq(double* v, int* ind, int* cls) {
  const double tol = F32C(1.e-2);
  auto rd = [](double a, double b) {
    if ((fabs(a) + fabs(b)) < F32C(1.0e-10)) return 0.0;
    return 2.0 * fabs(a - b) / (fabs(a) + fabs(b));
  };
  int s[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) s[i] += (rd(v[i], v[j]) < tol) ? 1 : 0;
  int s5 = s[0] + s[1] + s[2] + s[3];
  int mx = max(max(s[0], s[1]), max(s[2], s[3]));
  if (s5 == 8 && mx == 3) {
    for (int i = 0; i < 4; i++) if (s[i] == 2) s[i] += 1;
    s5 = s[0] + s[1] + s[2] + s[3];
  }
  int mn = min(min(s[0], s[1]), min(s[2], s[3]));
  if (s5 == 10 && mn == 2) { s[0] = s[1] = s[2] = s[3] = 4; s5 = 16; }
  if (s5 == 12 || s5 == 14) s5 = 16;
  for (int i = 0; i < 4; i++) ind[i] = i;
  auto swp = [&](int a, int b) {
    double t = v[a]; v[a] = v[b]; v[b] = t;
    int u = ind[a]; ind[a] = ind[b]; ind[b] = u;
    u = s[a]; s[a] = s[b]; s[b] = u;
  };
  if (s5 == 10) {
    for (int i = 0; i < 3; i++) if (s[i] == 1) swp(i, 3);
  } else if (s5 == 8) {
    for (int i = 2; i < 4; i++) if (rd(v[i], v[0]) < rd(v[1], v[0])) swp(i, 1);
  } else if (s5 == 6) {
    int j = 0;
    for (int i = 0; i < 4; i++) if (s[i] == 2) { swp(j, i); j++; }
  } else if (s5 != 4 && s5 != 16) {
    return false;
  }
  *cls = s5;
  return true;
}

// Delta = e_f - e_e at the small-tetrahedron vertices cr (linear in the
// big-tetrahedron corner values db); returns the Taylor threshold 10 max|Delta|
__device__ __forceinline__ double vertex_delta(const double cr[4][3], const double* db, double* ds) {
  double mx = 0.0;
  for (int i = 0; i < 4; i++) {
    ds[i] = (db[1] - db[0]) * cr[i][0] + (db[2] - db[0]) * cr[i][1] + (db[3] - db[0]) * cr[i][2] + db[0];
    if (fabs(ds[i]) > mx) mx = fabs(ds[i]);
  }
  return 10.0 * mx;
}

// One small tetrahedron (generictetra_par, for_tetrahedra3.f90:220-405),
// evaluated by the four lanes of an aligned lane group together, lane c
// handling big-tetrahedron corner c: lane c evaluates the closed form of
// small vertex j = c (each vertex once, as the reference does) with the
// atan / log of the vertex energies shared through the group (lane i
// evaluates the i-th sorted energy, or in classes 8/10 lanes 0 and 1 the
// two node energies), then adds sum_j w_j beta_c(j) (j ascending, beta = 0
// skipped) times vol to out[w] for frequencies j = c0+w in [jlo, jhi),
// evaluated at omega[j-shift].  The four lanes must run it together.
__device__ void small_tet(const double cr[4][3], const double* db, int corner, const double* __restrict__ omega,
                          int c0, int jlo, int jhi, int shift, int compat, double* out, int* nwarn, int* status) {
  double vec[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) vec[i][j] = cr[i + 1][j] - cr[0][j];
  double vol = 0.0;
  for (int i = 0; i < 3; i++) {
    int j = (i + 1) % 3, k = (j + 1) % 3;
    vol = vol + vec[0][i] * (vec[1][j] * vec[2][k] - vec[1][k] * vec[2][j]);
  }
  vol = fabs(vol);
  if (vol < F32C(1.e-10)) return;
  double ds[4], dsort[4];
  const double thr = vertex_delta(cr, db, ds);
  for (int i = 0; i < 4; i++) dsort[i] = ds[i];
  int ind[4], cls;
  if (!sort_eq(dsort, ind, &cls)) { *status |= ST_GEOM; return; }
  double beta[4];  // barycentric weight of this lane's corner at each small vertex
  for (int j = 0; j < 4; j++) beta[j] = (corner == 0) ? (1.0 - cr[j][0] - cr[j][1] - cr[j][2]) : cr[j][corner - 1];
  const int jv = corner;  // the small vertex this lane evaluates
  int v = 0;
  for (int i = 0; i < 4; i++) if (ind[i] == jv) v = i;
  VCoef c;
  vertex_coef(c, cls, dsort, v, compat);
  double s1, n03;
  taylor_coef(ds, jv, s1, n03);
  // this lane's share of the atan / log evaluations
  const double xsh = (cls == 8 || cls == 10) ? ((corner & 1) ? (cls == 8 ? c.ev[2] : c.ev[3]) : c.ev[0])
                                              : (corner == 0 ? dsort[0] : corner == 1 ? dsort[1]
                                                 : corner == 2 ? dsort[2] : dsort[3]);
  const bool f7b = compat & COMPAT_F7B;
  double tl[OMC];
#pragma unroll
  for (int w = 0; w < OMC; w++) tl[w] = 0.0;
#pragma unroll 1
  for (int w = 0; w < OMC; w++) {
    const int jj = c0 + w;
    double t0 = tl[0];
    if (jj >= jlo && jj < jhi) {
      const double om = omega[jj - shift];
      double wv;
      if (om > thr) {
        const double o2 = om * om;
        wv = s1 / (6.0e1 * o2) + n03 / (4.2e2 * (o2 * o2));
      } else if (cls == 16) {
        wv = check_weight(eval_cls16(c.ev[0], om), om, nwarn);
      } else {
        double tm, lm, T[4], L[4];
        tan_log(xsh, om, tm, lm);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          T[i] = __shfl(tm, i, 4);
          L[i] = __shfl(lm, i, 4);
        }
        if (cls == 4) {
          double Tv[4], Lv[4];
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const int r = f7b ? ((i - v + 4) & 3) : ((i + v) & 3);
            Tv[i] = r == 0 ? T[0] : r == 1 ? T[1] : r == 2 ? T[2] : T[3];
            Lv[i] = r == 0 ? L[0] : r == 1 ? L[1] : r == 2 ? L[2] : L[3];
          }
          wv = eval_cls4(c.t, c.w, c.cc, Tv, Lv, om);
        } else if (cls == 6) {
          double T0, L0;
          if (v < 2) {
            T0 = v == 0 ? T[0] : T[1];
            L0 = v == 0 ? L[0] : L[1];
          } else {
            tan_log(c.ev[0], om, T0, L0);  // the pair average
          }
          const bool sw = v == 3;  // vertex 4 takes d4 as its third energy
          wv = eval_cls6(c, om, T0, L0, sw ? T[3] : T[2], sw ? L[3] : L[2], sw ? T[2] : T[3], sw ? L[2] : L[3]);
        } else {
          wv = eval_kind2(c, om, T[0], L[0], T[1], L[1]);
        }
        wv = check_weight(wv, om, nwarn);
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const double wj = __shfl(wv, j, 4);
        if (beta[j] != 0.0) t0 = t0 + wj * beta[j];
      }
    }
#pragma unroll
    for (int i = 0; i < OMC - 1; i++) tl[i] = tl[i + 1];
    tl[OMC - 1] = t0;
  }
#pragma unroll
  for (int w = 0; w < OMC; w++) out[w] = out[w] + tl[w] * vol;
}

// Corner weight of one (occupied e, empty f) tetrahedron pair for corner
// `corner`, frequencies c0..c0+OMC-1 (convw1t_par); res[w] overwritten.
__device__ __noinline__ void corner_weight(const double* e, const double* f, double ef, int corner,
                                           const double* __restrict__ omega, int nom, int c0, int compat,
                                           double* res, int* nwarn, int* status) {
  for (int w = 0; w < OMC; w++) res[w] = 0.0;
  Nodes g;
  surf_nodes(g, e, f, ef);
  unrep_nodes(g);
  if (!rel_nodes(g, e, f, ef)) { *status |= ST_GEOM; return; }
  int8_t pos[MAXN][2];
  if (!sort_nodes(g, pos, status)) { *status |= ST_GEOM; return; }
  double db[4];
  for (int i = 0; i < 4; i++) db[i] = f[i] - e[i];
  const int n = g.n;
  if (n < 4) return;
  if (n > 8) { *status |= ST_GEOM; return; }
  double cr[6][3];
  if (n == 4) {
    for (int i = 0; i < 4; i++) for (int c = 0; c < 3; c++) cr[i][c] = g.x[i][c];
    small_tet(cr, db, corner, omega, c0, 0, nom, 0, compat, res, nwarn, status);
    return;
  }
  if (n == 5 || n == 6) {
    for (int i = 0; i < n; i++) for (int c = 0; c < 3; c++) cr[pos[i][0] - 1][c] = g.x[i][c];
    for (int s = 0; s + 4 <= n; s++)
      small_tet((const double(*)[3])cr[s], db, corner, omega, c0, 0, nom, 0, compat, res, nwarn, status);
    return;
  }
  // n = 7 (two pentahedra) or 8 (two prisms)
  const int nt = (n == 7) ? 2 : 3;
  double sub1[OMC], sub2[OMC], sub2s[OMC];
  for (int w = 0; w < OMC; w++) { sub1[w] = 0.0; sub2[w] = 0.0; sub2s[w] = 0.0; }
  const bool n78 = compat & COMPAT_N78;
  for (int s = 0; s < 2; s++) {
    for (int i = 0; i < n; i++)
      if (pos[i][s] != 0) for (int c = 0; c < 3; c++) cr[pos[i][s] - 1][c] = g.x[i][c];
    for (int t = 0; t < nt; t++) {
      if (s == 0) {
        small_tet((const double(*)[3])cr[t], db, corner, omega, c0, 0, n78 ? min(2, nom) : nom, 0, compat, sub1, nwarn,
                  status);
      } else {
        small_tet((const double(*)[3])cr[t], db, corner, omega, c0, 0, nom, 0, compat, sub2, nwarn, status);
        if (n78) small_tet((const double(*)[3])cr[t], db, corner, omega, c0, 2, nom, 2, compat, sub2s, nwarn, status);
      }
    }
  }
  for (int w = 0; w < OMC; w++) {
    int jj = c0 + w;
    double first = (!n78 || jj < 2) ? sub1[w] : sub2s[w];
    res[w] = first + sub2[w];
  }
}

// Insulating pair (every e below and every f above E_F by more than 1e-9):
// surfnodes..relnodes keep exactly the four corners (0,0,0), (0,0,1),
// (0,1,0), (1,0,0) -- no plane crosses the tetrahedron, and a margin of
// 1e-9 is far beyond the rounding of the intersection quotients -- so the
// polyhedron is the whole tetrahedron (nnod = 4): corner_weight reduces to
// small_tet on the unit corners, which full_tet_sym and pygw_tw_whole
// evaluate directly.
__device__ __forceinline__ bool insulating(const double* e, const double* f, double ef) {
  const double m = 1.0e-9;
  return fmax(fmax(e[0], e[1]), fmax(e[2], e[3])) < ef - m && fmin(fmin(f[0], f[1]), fmin(f[2], f[3])) > ef + m;
}

// Whole tetrahedron of class 16, 8 or 10 (all four, two pairs of, or three
// of the corner Delta's equal within sorteq's tolerance), all four corners at
// once.  The small tetrahedron is the big one (volume 1, barycentric weight 1
// at exactly one vertex per corner), so corner c's weight at a frequency is
// the closed form of its vertex alone.  In these classes the vertices share
// their energies (averages of the sorted Delta's) and fall into at most two
// coefficient sets -- sorted vertices below nA use set a, the others set b --
// so one frequency costs at most two atan / log evaluations for all four
// corners.  out + c*sc + w*sw receives corner c, frequency w.
__device__ __forceinline__ void full_tet_sym(int cls, const double* ds, const double* dsort, const int* ind,
                                             double thr, const double* __restrict__ omega, int nom, int compat,
                                             double* out, size_t sc, size_t sw, int* nwarn) {
  VCoef ca, cb;
  int nA;
  if (cls == 8) {
    vertex_coef(ca, 8, dsort, 0, compat);
    vertex_coef(cb, 8, dsort, 2, compat);
    nA = 2;
  } else if (cls == 10) {
    vertex_coef(ca, 10, dsort, 0, compat);
    vertex_coef(cb, 10, dsort, 3, compat);
    nA = 3;
  } else {
    vertex_coef(ca, 16, dsort, 0, compat);
    cb = ca;
    nA = 4;
  }
  // the two node energies of the two-node form (na = 0; nb = 2 for class 8, 3 for class 10)
  const double xa = ca.ev[0];
  const double xb = cls == 8 ? ca.ev[2] : ca.ev[3];
  double s1j[4], n03j[4], s1[4], n03[4];
  size_t oc[4];
  bool inA[4];
#pragma unroll
  for (int j = 0; j < 4; j++) taylor_coef(ds, j, s1j[j], n03j[j]);
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int j = ind[v];
    s1[v] = j == 0 ? s1j[0] : j == 1 ? s1j[1] : j == 2 ? s1j[2] : s1j[3];
    n03[v] = j == 0 ? n03j[0] : j == 1 ? n03j[1] : j == 2 ? n03j[2] : n03j[3];
    oc[v] = (size_t)(j == 0 ? 0 : 4 - j) * sc;   // corner of original vertex j
    inA[v] = v < nA;
  }
#pragma unroll 1
  for (int iw = 0; iw < nom; iw++) {
    const double om = omega[iw];
    double* o = out + (size_t)iw * sw;
    if (om > thr) {
      const double o2 = om * om;
#pragma unroll
      for (int v = 0; v < 4; v++) o[oc[v]] = s1[v] / (6.0e1 * o2) + n03[v] / (4.2e2 * (o2 * o2));
    } else if (cls == 16) {
      const double r = eval_cls16(xa, om);
#pragma unroll
      for (int v = 0; v < 4; v++) o[oc[v]] = check_weight(r, om, nwarn);
    } else {
      double Ta, La, Tb, Lb;
      tan_log(xa, om, Ta, La);
      tan_log(xb, om, Tb, Lb);
      const double ra = eval_kind2(ca, om, Ta, La, Tb, Lb);
      const double rb = eval_kind2(cb, om, Ta, La, Tb, Lb);
#pragma unroll
      for (int v = 0; v < 4; v++) o[oc[v]] = check_weight(inA[v] ? ra : rb, om, nwarn);
    }
  }
}

}  // namespace tw

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// incidence list: inc[k*24 + r] = 4*tet + corner, increasing in (tet, corner);
// cnt[k] = number found (24 on any mesh built by KQmesh.tetra).  One
// wavefront per k scans 64 tetrahedra at a time; a ballot keeps the order.
__global__ void __launch_bounds__(64) pygw_tw_incidence(const int* __restrict__ tetc, int ntet, int nk,
                                                        int* __restrict__ inc, int* __restrict__ cnt, int maxinc) {
  const int k = blockIdx.x;
  const int lane = threadIdx.x;
  if (k >= nk) return;
  int c = 0;
  bool bad = false;
  for (int t0 = 0; t0 < ntet; t0 += 64) {
    const int t = t0 + lane;
    int hit = 0;  // corners of tetrahedron t equal to k (bit i)
    if (t < ntet)
      for (int i = 0; i < 4; i++) {
        const int v = tetc[4 * t + i];
        bad |= (v < 0 || v >= nk);
        hit |= (v == k) << i;
      }
    // exclusive prefix of the hit counts over the wavefront
    const int nh = __popc(hit);
    int x = nh;
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    int slot = c + x - nh;
    for (int i = 0; i < 4; i++)
      if ((hit >> i) & 1) {
        if (slot < maxinc) inc[k * maxinc + slot] = 4 * t + i;
        slot++;
      }
    c += __shfl(x, 63, 64);
  }
  const bool anybad = __ballot(bad) != 0;
  if (lane == 0) cnt[k] = anybad ? maxinc + 1 : c;  // > maxinc: malformed tetc, nothing is read
}

// Per chunk of band pairs (pair index q in the output layout's order:
// layout 0 q = i_o + n_o*i_e, layout 1 q = i_o*n_e + i_e; buffers at pitch
// pp >= the chunk's pcn pairs):
//   pygw_tw_table    one lane per (k, pair): Delta and its atan / log over
//                    all frequencies;
//   pygw_tw_tet      one lane per (tetrahedron, pair): the four corner
//                    weights over all frequencies into
//                    w1t[T][corner][w][q - q0] (convw1t_par's w1t(4, nom))
//                    for whole tetrahedra of classes 16, 8, 10, zeros where
//                    no band crosses E_F in T; whole tetrahedra of classes
//                    4 and 6 go to one list, polyhedra cut by E_F (and
//                    classification errors) to another;
//   pygw_tw_whole    one lane per (listed pair, corner), classes 4 and 6
//                    from the table;
//   pygw_tw_general  four lanes per listed pair (one per corner, running the
//                    general polyhedron decomposition together);
//   pygw_tw_sum      one lane per (k, frequency, pair): kcw = sum over the
//                    incidences of k in increasing tetrahedron order of
//                    w1t * 6/N_tet, the scatter order of
//                    for_tetrahedra.f90:445-456.
__device__ __forceinline__ void tw_pair(int q, int layout, int no, int ne, int& ib, int& jb) {
  if (layout == 0) { ib = q % no; jb = q / no; }
  else { ib = q / ne; jb = q % ne; }
}

// band energies at the corners of T (occupied band ib) and of its k-q
// partner tl (empty band jba); false (nothing read) for a corner index
// outside the k-mesh
__device__ __forceinline__ bool tw_corners(const double* __restrict__ enk, int nbnd, int nk,
                                           const int* __restrict__ tetc, int T, int tl, int ib, int jba, double* e,
                                           double* f) {
  int a[4], b[4];
  bool ok = true;
  for (int i = 0; i < 4; i++) {
    a[i] = tetc[4 * T + i];
    b[i] = tetc[4 * tl + i];
    ok &= a[i] >= 0 && a[i] < nk && b[i] >= 0 && b[i] < nk;
  }
  if (!ok) return false;
  for (int i = 0; i < 4; i++) {
    f[i] = enk[jba + (size_t)nbnd * b[i]];
    e[i] = enk[ib + (size_t)nbnd * a[i]];
  }
  return true;
}

// atan / log table of one chunk: for k-point k and pair q, Delta =
// e_f(kqid[k]) - e_e(k) (a corner value of every whole tetrahedron with
// corner k whose k-q partner corner is kqid[k]), dtab[k][q] = Delta and
// tab[k][w][q] = (atan(Delta/w), log(Delta^2 + w^2)) as in the closed forms
__global__ void __launch_bounds__(64) pygw_tw_table(const double* __restrict__ enk, int nbnd,
                                                    const int* __restrict__ kqid, int ncg, int numin, int no, int ne,
                                                    const double* __restrict__ omega, int nom, int nk, int q0, int pcn,
                                                    int pp, int layout, double2* __restrict__ tab,
                                                    double* __restrict__ dtab) {
  const int ql = blockIdx.x * 64 + threadIdx.x;
  const int k = blockIdx.y;
  if (ql >= pcn || k >= nk) return;
  int ib, jb;
  tw_pair(q0 + ql, layout, no, ne, ib, jb);
  const int kk = kqid[k];
  if (kk < 0 || kk >= nk) {
    dtab[(size_t)k * pp + ql] = __longlong_as_double(0x7ff8000000000001LL);  // matches no corner value
    return;
  }
  const double d = enk[jb + ncg + numin + (size_t)nbnd * kk] - enk[ib + (size_t)nbnd * k];
  dtab[(size_t)k * pp + ql] = d;
  double2* o = tab + (size_t)k * nom * pp + ql;
  for (int w = 0; w < nom; w++) {
    double t, l;
    tw::tan_log(d, omega[w], t, l);
    o[(size_t)w * pp] = double2{t, l};
  }
}

// append (T, ql) of the flagged lanes to a list (one atomic per wavefront)
__device__ __forceinline__ void tw_append(bool flag, int2* __restrict__ list, int* __restrict__ count, int T, int ql) {
  const unsigned long long m = __ballot(flag);
  if (!m) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, leader, 64);
  if (flag) list[base + __popcll(m & ((1ull << lane) - 1ull))] = make_int2(T, ql);
}

__global__ void __launch_bounds__(64) pygw_tw_tet(
    const double* __restrict__ enk, int nbnd, double ef, const int* __restrict__ kqid, int ncg, int numin, int no,
    int ne, const double* __restrict__ omega, int nom, const int* __restrict__ tetc, int nk, int ntet, int q0, int pcn,
    int pp, int compat, int layout, const double* __restrict__ dtab, double* __restrict__ w1t, int2* __restrict__ wlist,
    int* __restrict__ wcount, int2* __restrict__ glist, int* __restrict__ gcount, int* __restrict__ dstat) {
  const int nbp = (pcn + 63) >> 6;
  const int T = blockIdx.x / nbp;
  const int ql = (blockIdx.x - T * nbp) * 64 + threadIdx.x;
  const bool act = ql < pcn && T < ntet;
  int nwarn = 0, status = 0;
  bool general = false, whole = false;
  if (act) {
    int ib, jb;
    tw_pair(q0 + ql, layout, no, ne, ib, jb);
    const size_t sc = (size_t)nom * pp, sw = (size_t)pp;
    double* out = w1t + (size_t)T * 4 * sc + ql;
    const int kl = kqid[T / 6];
    int cat = 0;  // 0: no weight, 1: whole tetrahedron, 2: polyhedron cut by E_F
    double e[4], f[4];
    if (kl < 0 || kl >= nk || !tw_corners(enk, nbnd, nk, tetc, T, 6 * kl + T % 6, ib, jb + ncg + numin, e, f)) {
      status |= 4;
    } else {
      const double fmx = fmax(fmax(f[0], f[1]), fmax(f[2], f[3]));
      const double emn = fmin(fmin(e[0], e[1]), fmin(e[2], e[3]));
      if (!(fmx <= ef || emn > ef)) cat = tw::insulating(e, f, ef) ? 1 : 2;
    }
    if (cat == 0) {
      for (int c = 0; c < 4; c++)
        for (int w = 0; w < nom; w++) out[c * sc + w * sw] = 0.0;
    } else if (cat == 1) {
      // the whole tetrahedron: vertex energies by the small-tetrahedron formula
      const double cr[4][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 1.0}, {0.0, 1.0, 0.0}, {1.0, 0.0, 0.0}};
      double db[4], ds[4], dsort[4];
      for (int i = 0; i < 4; i++) db[i] = f[i] - e[i];
      const double thr = tw::vertex_delta(cr, db, ds);
      for (int i = 0; i < 4; i++) dsort[i] = ds[i];
      int ind[4], cls;
      if (!tw::sort_eq(dsort, ind, &cls)) {
        general = true;  // the general path reports the classification error
      } else if (cls == 16 || cls == 8 || cls == 10) {
        tw::full_tet_sym(cls, ds, dsort, ind, thr, omega, nom, compat, out, sc, sw, &nwarn);
      } else {
        // classes 4 and 6: per corner with the tabulated atan / log when every
        // vertex energy is its corner's table entry bit for bit
        bool hit = true;
        for (int j = 0; j < 4; j++) {  // vertex j sits at big corner 0, 3, 2, 1
          const int kc = tetc[4 * T + (j == 0 ? 0 : 4 - j)];
          hit &= __double_as_longlong(ds[j]) == __double_as_longlong(dtab[(size_t)kc * pp + ql]);
        }
        if (hit) whole = true;
        else general = true;
      }
    } else {
      general = true;
    }
  }
  tw_append(whole, wlist, wcount, T, ql);
  tw_append(general, glist, gcount, T, ql);
  if (status) atomicOr(&dstat[0], status);
  if (nwarn) atomicAdd(&dstat[1], nwarn);
}

// Whole tetrahedra of class 4 or 6 listed by pygw_tw_tet, one lane per
// (listed pair, corner): the closed form of the corner's vertex with the
// atan / log of the sorted vertex energies read from the table (in class 6
// the vertices 3, 4 use the average of the equal pair, evaluated here).
__global__ void __launch_bounds__(64) pygw_tw_whole(
    const double* __restrict__ enk, int nbnd, const int* __restrict__ kqid, int ncg, int numin, int no, int ne,
    const double* __restrict__ omega, int nom, const int* __restrict__ tetc, int nk, int q0, int pp, int compat,
    int layout, const double2* __restrict__ tab, const int2* __restrict__ wlist, const int* __restrict__ wcount,
    double* __restrict__ w1t, int* __restrict__ dstat) {
  const int n = 4 * wcount[0];
  int nwarn = 0;
  const size_t sc = (size_t)nom * pp, sw = (size_t)pp;
  for (int u = blockIdx.x * 64 + threadIdx.x; u < n; u += gridDim.x * 64) {
    const int2 it = wlist[u >> 2];
    const int corner = u & 3, T = it.x, ql = it.y;
    int ib, jb;
    tw_pair(q0 + ql, layout, no, ne, ib, jb);
    double e[4], f[4];
    tw_corners(enk, nbnd, nk, tetc, T, 6 * kqid[T / 6] + T % 6, ib, jb + ncg + numin, e, f);  // checked in pass 1
    const double cr[4][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 1.0}, {0.0, 1.0, 0.0}, {1.0, 0.0, 0.0}};
    double db[4], ds[4], dsort[4];
    for (int i = 0; i < 4; i++) db[i] = f[i] - e[i];
    const double thr = tw::vertex_delta(cr, db, ds);
    for (int i = 0; i < 4; i++) dsort[i] = ds[i];
    int ind[4], cls;
    tw::sort_eq(dsort, ind, &cls);  // class 4 or 6, as in pass 1
    const int j = corner == 0 ? 0 : 4 - corner;  // original vertex of this corner
    int v = 0;
    for (int i = 0; i < 4; i++) if (ind[i] == j) v = i;
    tw::VCoef c;
    tw::vertex_coef(c, cls, dsort, v, compat);
    double s1, n03;
    tw::taylor_coef(ds, j, s1, n03);
    // table rows of the sorted energies, then of the vertex energies ev[i]
    const double2* row[4];
    for (int i = 0; i < 4; i++) {
      const int ji = ind[i];
      row[i] = tab + (size_t)tetc[4 * T + (ji == 0 ? 0 : 4 - ji)] * nom * pp + ql;
    }
    const double2* pe[4];
    if (cls == 4) {
      for (int i = 0; i < 4; i++) pe[i] = row[(compat & tw::COMPAT_F7B) ? ((i - v + 4) & 3) : ((i + v) & 3)];
    } else {
      pe[0] = row[v < 2 ? v : 0];   // v >= 2: ev[0] is the pair average (computed below)
      pe[1] = pe[0];
      pe[2] = row[v == 3 ? 3 : 2];
      pe[3] = row[v == 3 ? 2 : 3];
    }
    double* out = w1t + ((size_t)T * 4 + corner) * sc + ql;
    for (int iw = 0; iw < nom; iw++) {
      const double om = omega[iw];
      double r;
      if (om > thr) {
        const double o2 = om * om;
        r = s1 / (6.0e1 * o2) + n03 / (4.2e2 * (o2 * o2));
      } else if (cls == 4) {
        double Tt[4], Lt[4];
        for (int i = 0; i < 4; i++) {
          const double2 x = pe[i][(size_t)iw * sw];
          Tt[i] = x.x;
          Lt[i] = x.y;
        }
        r = tw::check_weight(tw::eval_cls4(c.t, c.w, c.cc, Tt, Lt, om), om, &nwarn);
      } else {
        double T0, L0;
        if (v < 2) {
          const double2 x = pe[0][(size_t)iw * sw];
          T0 = x.x;
          L0 = x.y;
        } else {
          tw::tan_log(c.ev[0], om, T0, L0);
        }
        const double2 x2 = pe[2][(size_t)iw * sw], x3 = pe[3][(size_t)iw * sw];
        r = tw::check_weight(tw::eval_cls6(c, om, T0, L0, x2.x, x2.y, x3.x, x3.y), om, &nwarn);
      }
      out[(size_t)iw * sw] = r;
    }
  }
  if (nwarn) atomicAdd(&dstat[1], nwarn);
}

__global__ void __launch_bounds__(64) pygw_tw_general(
    const double* __restrict__ enk, int nbnd, double ef, const int* __restrict__ kqid, int ncg, int numin, int no,
    int ne, const double* __restrict__ omega, int nom, const int* __restrict__ tetc, int nk, int q0, int pp,
    int compat, int layout, const int2* __restrict__ glist, const int* __restrict__ gcount, double* __restrict__ w1t,
    int* __restrict__ dstat) {
  const int n = 4 * gcount[0];
  int nwarn = 0, status = 0;
  const size_t sc = (size_t)nom * pp, sw = (size_t)pp;
  for (int u = blockIdx.x * 64 + threadIdx.x; u < n; u += gridDim.x * 64) {
    const int2 it = glist[u >> 2];
    const int corner = u & 3, T = it.x, ql = it.y;
    int ib, jb;
    tw_pair(q0 + ql, layout, no, ne, ib, jb);
    double e[4], f[4];
    tw_corners(enk, nbnd, nk, tetc, T, 6 * kqid[T / 6] + T % 6, ib, jb + ncg + numin, e, f);  // checked in pass 1
    double* out = w1t + ((size_t)T * 4 + corner) * sc + ql;
    for (int c0 = 0; c0 < nom; c0 += tw::OMC) {
      double res[tw::OMC];
      tw::corner_weight(e, f, ef, corner, omega, nom, c0, compat, res, &nwarn, &status);
      for (int i = 0; i < tw::OMC && c0 + i < nom; i++) out[(size_t)(c0 + i) * sw] = res[i];
    }
  }
  if (status) atomicOr(&dstat[0], status);
  if (nwarn) atomicAdd(&dstat[1], nwarn);
}

// layout 0: reference F-order kcw[n_o, n_e, N_k, nom]
// layout 1: kernel-native kcw[N_k][nom][n_o*n_e], pair p = i_o*n_e + i_e
// one lane per (k, frequency, pair): the 24 incidences summed in order
__global__ void __launch_bounds__(64) pygw_tw_sum(const double* __restrict__ w1t, const int* __restrict__ inc,
                                                  const int* __restrict__ cnt, int maxinc, int nk, int ntet, int nom,
                                                  int q0, int pcn, int pp, int np, int layout,
                                                  double* __restrict__ kcw,
                                                  int* __restrict__ dstat) {
  const int ql = blockIdx.x * 64 + threadIdx.x;
  const int w = blockIdx.y, k = blockIdx.z;
  if (ql >= pcn || k >= nk || w >= nom) return;
  const double wtet = 1.0 / ntet;
  int ninc = cnt[k];
  if (ninc > maxinc) {
    ninc = 0;
    if (ql == 0 && w == 0) atomicOr(&dstat[0], 4);
  }
  const size_t sc = (size_t)nom * pp;
  const double* src = w1t + (size_t)w * pp + ql;
  double acc = 0.0;
  for (int r = 0; r < ninc; r++) acc = acc + src[(size_t)inc[k * maxinc + r] * sc] * 6 * wtet;
  const size_t q = (size_t)(q0 + ql);
  const size_t off = layout == 0 ? q + (size_t)np * ((size_t)k + (size_t)nk * w)
                                 : q + (size_t)np * ((size_t)w + (size_t)nom * k);
  kcw[off] = acc;
}

// SVD-basis weights (bz_calcqdepw_par3, for_tetrahedra.f90:466-542):
//   kcw_l[k][l][p] = sum_w kcw[k][w][p] * dUl[w][l]      (dUl Fortran (nom, nl))
// The reference projects each tetrahedron's w1t before the scatter
// (w2t = matmul(w1t, dUl), :532); the projection is linear, so projecting
// the gathered kcw gives the same sums up to rounding.
// layout 0: out Fortran [n_o, n_e, N_k, nl]; layout 1: out [N_k][nl][n_o*n_e].
__global__ void __launch_bounds__(256) pygw_tw_project(const double* __restrict__ kcw, int np, int nom, int nk,
                                                       const double* __restrict__ dUl, int nl, int layout, int no,
                                                       int ne, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (t >= np || k >= nk) return;
  // thread t holds pair index in the output layout's order
  int p;
  if (layout == 0) { const int ib = t % no, jb = t / no; p = ib * ne + jb; }
  else p = t;
  const double* src = kcw + (size_t)k * nom * np + p;
  for (int l = 0; l < nl; l++) {
    double s = 0.0;
    for (int w = 0; w < nom; w++) s = s + src[(size_t)w * np] * dUl[w + (size_t)nom * l];
    size_t off = (layout == 0) ? (size_t)t + (size_t)np * ((size_t)k + (size_t)nk * l)
                               : (size_t)t + (size_t)np * ((size_t)l + (size_t)nl * k);
    out[off] = s;
  }
}
