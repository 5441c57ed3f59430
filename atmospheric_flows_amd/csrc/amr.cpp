// Locally refined meshes: geometry, face table and multigrid hierarchy
// (see amr.h).
#include "amr.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <tuple>
#include <unordered_map>

namespace atmo {

namespace {

// (tree, level, sx, sz) packed into one key
inline unsigned long long key_of(long tree, int l, int sx, int sz) {
  return ((unsigned long long)tree << 40) | ((unsigned long long)l << 32) | ((unsigned long long)sx << 16) |
         (unsigned long long)sz;
}

struct Index {
  std::unordered_map<unsigned long long, int> map;
  int find(long tree, int l, int sx, int sz) const {
    if (l < 0) return -1;
    auto it = map.find(key_of(tree, l, sx, sz));
    return it == map.end() ? -1 : it->second;
  }
};

void geometry(const MeshDesc& b, const std::vector<QCell>& cells, std::vector<double>& cgeo) {
  cgeo.resize(6 * cells.size());
  for (size_t c = 0; c < cells.size(); ++c) {
    const QCell& q = cells[c];
    const double n = (double)(1 << q.l);
    const double xl = b.Lx * q.bi / b.nx, xr = b.Lx * (q.bi + 1) / b.nx;
    const double hl = b.h[q.bi], hr = b.h[q.bi + 1];
    const double zb[4] = {hl + (b.H - hl) * q.bj / b.nz, hr + (b.H - hr) * q.bj / b.nz,
                          hl + (b.H - hl) * (q.bj + 1) / b.nz, hr + (b.H - hr) * (q.bj + 1) / b.nz};
    auto zmap = [&](double xi, double eta) {
      return (1 - xi) * (1 - eta) * zb[0] + xi * (1 - eta) * zb[1] + (1 - xi) * eta * zb[2] + xi * eta * zb[3];
    };
    const double x0 = q.sx / n, x1 = (q.sx + 1) / n, e0 = q.sz / n, e1 = (q.sz + 1) / n;
    double* g = &cgeo[6 * c];
    g[0] = xl + x0 * (xr - xl);
    g[1] = (xl + x1 * (xr - xl)) - g[0];
    g[2] = zmap(x0, e0);
    g[3] = zmap(x1, e0);
    g[4] = zmap(x0, e1);
    g[5] = zmap(x1, e1);
  }
}

// Face table of the finest level (see MeshDev): neighbour of the same size,
// of the next larger size (this face is one half of its face) or split
// between two smaller cells.
void faces(const MeshDesc& b, const std::vector<QCell>& cells, const Index& idx, std::vector<int>& cface) {
  cface.assign(16 * cells.size(), 0);
  for (size_t c = 0; c < cells.size(); ++c) {
    const QCell& q = cells[c];
    const int n = 1 << q.l;
    for (int f = 0; f < 4; ++f) {
      int* r = &cface[16 * c + 4 * f];
      int bi = q.bi, bj = q.bj, sx = q.sx, sz = q.sz;
      switch (f) {
        case 0: if (--sx < 0) { sx = n - 1; --bi; } break;
        case 1: if (++sx == n) { sx = 0; ++bi; } break;
        case 2: if (--sz < 0) { sz = n - 1; --bj; } break;
        default: if (++sz == n) { sz = 0; ++bj; } break;
      }
      if (bj < 0 || bj >= b.nz || ((bi < 0 || bi >= b.nx) && !b.periodic)) {
        r[0] = FACE_BDRY;
        continue;
      }
      bi = (bi + b.nx) % b.nx;
      const long tree = (long)bj * b.nx + bi;
      int nb = idx.find(tree, q.l, sx, sz);
      if (nb >= 0) {
        r[0] = FACE_SAME;
        r[1] = nb;
        continue;
      }
      nb = idx.find(tree, q.l - 1, sx >> 1, sz >> 1);
      if (nb >= 0) {
        r[0] = FACE_COARSER;
        r[1] = nb;
        r[3] = f < 2 ? (q.sz & 1) : (q.sx & 1);
        continue;
      }
      // split face: the two cells of the next level along it, lower/left half first
      int h[2];
      for (int a = 0; a < 2; ++a) {
        const int cx = f < 2 ? 2 * sx + (f == 0 ? 1 : 0) : 2 * sx + a;
        const int cz = f < 2 ? 2 * sz + a : 2 * sz + (f == 2 ? 1 : 0);
        h[a] = idx.find(tree, q.l + 1, cx, cz);
        if (h[a] < 0) throw std::invalid_argument("locally refined mesh: not a 2:1-balanced cover of the base mesh");
      }
      r[0] = FACE_FINER;
      r[1] = h[0];
      r[2] = h[1];
    }
  }
}

}  // namespace

std::vector<AmrLevelHost> build_amr_levels(const MeshDesc& b, const std::vector<QCell>& leaves) {
  if ((int)b.h.size() != b.nx + 1) throw std::invalid_argument("locally refined mesh: need nx+1 terrain heights");
  std::vector<AmrLevelHost> out;
  std::vector<QCell> cur = leaves;
  // the leaves: validate, index, geometry, faces; area coverage check
  {
    Index idx;
    double area = 0.0;
    for (size_t c = 0; c < cur.size(); ++c) {
      const QCell& q = cur[c];
      if (q.bi < 0 || q.bi >= b.nx || q.bj < 0 || q.bj >= b.nz || q.l < 0 || q.l > 12 || q.sx < 0 || q.sz < 0 ||
          q.sx >= (1 << q.l) || q.sz >= (1 << q.l))
        throw std::invalid_argument("locally refined mesh: cell out of range");
      if (c > 0 && (long)cur[c - 1].bj * b.nx + cur[c - 1].bi > (long)q.bj * b.nx + q.bi)
        throw std::invalid_argument("locally refined mesh: leaves not in base-cell order");
      if (!idx.map.emplace(key_of((long)q.bj * b.nx + q.bi, q.l, q.sx, q.sz), (int)c).second)
        throw std::invalid_argument("locally refined mesh: duplicate cell");
      area += std::ldexp(1.0, -2 * q.l);
    }
    if (std::abs(area - (double)b.nx * b.nz) > 1e-9)
      throw std::invalid_argument("locally refined mesh: leaves do not cover the base mesh");
    AmrLevelHost L;
    L.ncell = (int)cur.size();
    L.cells = cur;
    geometry(b, cur, L.cgeo);
    faces(b, cur, idx, L.cface);
    out.push_back(std::move(L));
  }
  // global coarsening down to the base mesh
  while (true) {
    bool refined = false;
    for (const QCell& q : cur) refined |= q.l > 0;
    if (!refined) break;
    std::map<std::tuple<int, int, int, int, int>, int> siblings;
    for (const QCell& q : cur)
      if (q.l > 0) ++siblings[std::make_tuple(q.bi, q.bj, q.l - 1, q.sx >> 1, q.sz >> 1)];
    std::vector<QCell> next;
    std::vector<int> child;
    std::vector<int> parent(2 * cur.size());
    std::map<std::tuple<int, int, int, int, int>, int> placed;
    for (size_t c = 0; c < cur.size(); ++c) {
      const QCell& q = cur[c];
      const auto pk = std::make_tuple(q.bi, q.bj, q.l - 1, q.sx >> 1, q.sz >> 1);
      if (q.l > 0 && siblings[pk] == 4) {
        auto it = placed.find(pk);
        int pc;
        if (it == placed.end()) {
          pc = (int)next.size();
          placed[pk] = pc;
          next.push_back({q.bi, q.bj, q.l - 1, q.sx >> 1, q.sz >> 1});
          child.insert(child.end(), {-1, -1, -1, -1});
        } else {
          pc = it->second;
        }
        const int code = (q.sx & 1) + 2 * (q.sz & 1);
        child[4 * pc + code] = (int)c;
        parent[2 * c] = pc;
        parent[2 * c + 1] = code;
      } else {
        const int pc = (int)next.size();
        next.push_back(q);
        child.insert(child.end(), {(int)c, -1, -1, -1});
        parent[2 * c] = pc;
        parent[2 * c + 1] = -1;
      }
    }
    out.back().parent = parent;
    AmrLevelHost L;
    L.ncell = (int)next.size();
    L.cells = next;
    L.child = child;
    geometry(b, next, L.cgeo);
    out.push_back(std::move(L));
    cur = next;
  }
  AmrLevelHost& base = out.back();
  if (base.ncell != b.nx * b.nz) throw std::invalid_argument("locally refined mesh: leaves do not cover the base mesh");
  for (int c = 0; c < base.ncell; ++c)
    if ((long)base.cells[c].bj * b.nx + base.cells[c].bi != c)
      throw std::invalid_argument("locally refined mesh: leaves not in base-cell order");
  base.base = true;
  return out;
}

double amr_min_cell_diameter(const AmrLevelHost& L) {
  double d = 1e300;
  for (int c = 0; c < L.ncell; ++c) {
    const double* g = &L.cgeo[6 * c];
    const double d1 = std::hypot(g[1], g[5] - g[2]), d2 = std::hypot(g[1], g[4] - g[3]);
    d = std::min(d, std::max(d1, d2));
  }
  return d;
}

}  // namespace atmo
