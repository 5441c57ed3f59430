// VTU output of (rho, u, p) in the form of the reference's output_results
// (Atmospheric_Flow.cc:886-918): DataOut with patches of `degree`
// subdivisions per cell (build_patches(mapping, degree_u)), point data
// "rho", "u" (a 3-component vector, zero third component in 2D) and "p",
// each cell's points its own (the fields are discontinuous), or, with
// high_order, one VTK Lagrange quadrilateral of (k+1)^2 equispaced nodes per
// cell (the reference's solution_high_order-*.vtu, write_higher_order_cells).
// Data arrays are VTK inline binary: base64 of a UInt32 byte count followed
// by the little-endian values.
#include "vtu.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "basis.h"

namespace atmo {

namespace {

const char* kB64 = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

void put_base64(FILE* f, const std::vector<uint8_t>& bytes) {
  std::string out;
  out.reserve(4 * ((bytes.size() + 2) / 3));
  size_t i = 0;
  for (; i + 2 < bytes.size(); i += 3) {
    const uint32_t v = (uint32_t)bytes[i] << 16 | (uint32_t)bytes[i + 1] << 8 | bytes[i + 2];
    out += kB64[v >> 18];
    out += kB64[(v >> 12) & 63];
    out += kB64[(v >> 6) & 63];
    out += kB64[v & 63];
  }
  if (i < bytes.size()) {
    const uint32_t v = (uint32_t)bytes[i] << 16 | (i + 1 < bytes.size() ? (uint32_t)bytes[i + 1] << 8 : 0u);
    out += kB64[v >> 18];
    out += kB64[(v >> 12) & 63];
    out += i + 1 < bytes.size() ? kB64[(v >> 6) & 63] : '=';
    out += '=';
  }
  std::fputs(out.c_str(), f);
}

template <typename T>
void data_array(FILE* f, const char* type, const char* name, int ncomp, const std::vector<T>& v) {
  std::fprintf(f, "        <DataArray type=\"%s\"", type);
  if (name) std::fprintf(f, " Name=\"%s\"", name);
  std::fprintf(f, " NumberOfComponents=\"%d\" format=\"binary\">\n", ncomp);
  const uint32_t nb = (uint32_t)(v.size() * sizeof(T));
  std::vector<uint8_t> bytes(4 + nb);
  for (int b = 0; b < 4; ++b) bytes[b] = (uint8_t)(nb >> (8 * b));
  if (nb) std::memcpy(bytes.data() + 4, v.data(), nb);
  put_base64(f, bytes);
  std::fputs("\n        </DataArray>\n", f);
}

// Lagrange basis on the k+1 Gauss-Lobatto nodes of FE_DGQ(k) at x in [0, 1]
template <int K>
void lagrange(double x, double (&l)[K + 1]) {
  constexpr basis::Tables<K> t = basis::make_tables<K>();
  for (int i = 0; i <= K; ++i) {
    double v = 1.0;
    for (int j = 0; j <= K; ++j)
      if (j != i) v *= (x - t.gll[j]) / (t.gll[i] - t.gll[j]);
    l[i] = v;
  }
}

// VTK Lagrange quadrilateral of order n: node (a, b) (a along xi, b along
// eta, 0..n) in VTK's point order (corners, then the edges bottom, right,
// top, left each by increasing coordinate, then the interior row by row)
int vtk_lagrange_index(int a, int b, int n) {
  const bool ea = a == 0 || a == n, eb = b == 0 || b == n;
  if (ea && eb) return a == 0 ? (b == 0 ? 0 : 3) : (b == 0 ? 1 : 2);
  if (eb) return 4 + (b == 0 ? 0 : 2 * (n - 1)) + (a - 1);
  if (ea) return 4 + (a == 0 ? 3 * (n - 1) : (n - 1)) + (b - 1);
  return 4 + 4 * (n - 1) + (a - 1) + (n - 1) * (b - 1);
}

template <int K>
void write(const MeshDesc& m, const std::vector<double>& cgeo, const double* rho, const double* u, const double* p,
           const std::string& path, bool high_order) {
  constexpr int ND = K + 1, ND2 = ND * ND;
  const int n = K;  // subdivisions per cell direction (DataOut::build_patches(mapping, degree))
  const long ncell = cgeo.empty() ? (long)m.nx * m.nz : (long)(cgeo.size() / 6);
  const long Ns = ncell * ND2;
  const int ppc = (n + 1) * (n + 1);  // points per mesh cell
  // the file first: an unwritable path is refused before any field value is read
  struct Closer {
    void operator()(FILE* fp) const { std::fclose(fp); }
  };
  std::unique_ptr<FILE, Closer> file(std::fopen(path.c_str(), "w"));
  if (!file) throw std::invalid_argument("write_vtu: cannot open " + path);
  FILE* f = file.get();
  std::vector<double> pts, vr, vu, vp;
  pts.reserve(3 * ncell * ppc);
  vr.reserve(ncell * ppc);
  vu.reserve(3 * ncell * ppc);
  vp.reserve(ncell * ppc);
  // per mesh cell, the points in output order: lexicographic for the patches, VTK's order for Lagrange cells
  std::vector<std::pair<int, int>> order(ppc);
  for (int b = 0; b <= n; ++b)
    for (int a = 0; a <= n; ++a) order[high_order ? vtk_lagrange_index(a, b, n) : b * (n + 1) + a] = {a, b};
  for (long c = 0; c < ncell; ++c) {
    double x0, dx, z00, z10, z01, z11;
    if (cgeo.empty()) {
      const int i = (int)(c % m.nx), j = (int)(c / m.nx);
      dx = m.Lx / m.nx;
      x0 = i * dx;
      const double f0 = (double)j / m.nz, f1 = (double)(j + 1) / m.nz;
      z00 = m.h[i] + (m.H - m.h[i]) * f0;
      z10 = m.h[i + 1] + (m.H - m.h[i + 1]) * f0;
      z01 = m.h[i] + (m.H - m.h[i]) * f1;
      z11 = m.h[i + 1] + (m.H - m.h[i + 1]) * f1;
    } else {
      const double* g = &cgeo[6 * c];
      x0 = g[0];
      dx = g[1];
      z00 = g[2];
      z10 = g[3];
      z01 = g[4];
      z11 = g[5];
    }
    for (const auto& ab : order) {
      const double xi = (double)ab.first / n, eta = (double)ab.second / n;
      pts.push_back(x0 + xi * dx);
      pts.push_back((1 - xi) * (1 - eta) * z00 + xi * (1 - eta) * z10 + (1 - xi) * eta * z01 + xi * eta * z11);
      pts.push_back(0.0);
      double l1[ND], l2[ND];
      lagrange<K>(xi, l1);
      lagrange<K>(eta, l2);
      double sr = 0, su0 = 0, su1 = 0, sp = 0;
      for (int jj = 0; jj < ND; ++jj)
        for (int ii = 0; ii < ND; ++ii) {
          const double w = l1[ii] * l2[jj];
          const long d = c * ND2 + jj * ND + ii;
          sr += w * rho[d];
          su0 += w * u[d];
          su1 += w * u[Ns + d];
          sp += w * p[d];
        }
      vr.push_back(sr);
      vu.push_back(su0);
      vu.push_back(su1);
      vu.push_back(0.0);
      vp.push_back(sp);
    }
  }
  std::vector<int32_t> conn, offs;
  std::vector<uint8_t> types;
  if (high_order) {
    for (long c = 0; c < ncell; ++c) {
      for (int q = 0; q < ppc; ++q) conn.push_back((int32_t)(c * ppc + q));
      offs.push_back((int32_t)((c + 1) * ppc));
      types.push_back(70);  // VTK_LAGRANGE_QUADRILATERAL
    }
  } else {
    for (long c = 0; c < ncell; ++c)
      for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a) {
          const int32_t base = (int32_t)(c * ppc);
          conn.push_back(base + b * (n + 1) + a);
          conn.push_back(base + b * (n + 1) + a + 1);
          conn.push_back(base + (b + 1) * (n + 1) + a + 1);
          conn.push_back(base + (b + 1) * (n + 1) + a);
          offs.push_back((int32_t)conn.size());
          types.push_back(9);  // VTK_QUAD
        }
  }
  std::fprintf(f, "<?xml version=\"1.0\"?>\n");
  std::fprintf(f, "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
                  "header_type=\"UInt32\">\n  <UnstructuredGrid>\n");
  std::fprintf(f, "    <Piece NumberOfPoints=\"%ld\" NumberOfCells=\"%zu\">\n", ncell * ppc, types.size());
  std::fprintf(f, "      <Points>\n");
  data_array(f, "Float64", nullptr, 3, pts);
  std::fprintf(f, "      </Points>\n      <Cells>\n");
  data_array(f, "Int32", "connectivity", 1, conn);
  data_array(f, "Int32", "offsets", 1, offs);
  data_array(f, "UInt8", "types", 1, types);
  std::fprintf(f, "      </Cells>\n      <PointData Scalars=\"rho\" Vectors=\"u\">\n");
  data_array(f, "Float64", "rho", 1, vr);
  data_array(f, "Float64", "u", 3, vu);
  data_array(f, "Float64", "p", 1, vp);
  std::fprintf(f, "      </PointData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
  const bool ok = std::ferror(f) == 0 && std::fclose(file.release()) == 0;
  if (!ok) throw std::runtime_error("write_vtu: write error on " + path);
}

}  // namespace

void write_vtu(int k, const MeshDesc& m, const std::vector<double>& cgeo, const double* rho, const double* u,
               const double* p, const std::string& path, bool high_order) {
  if (k == 1) write<1>(m, cgeo, rho, u, p, path, high_order);
  else if (k == 2) write<2>(m, cgeo, rho, u, p, path, high_order);
  else if (k == 3) write<3>(m, cgeo, rho, u, p, path, high_order);
  else throw std::invalid_argument("write_vtu: degree must be 1..3");
}

}  // namespace atmo
