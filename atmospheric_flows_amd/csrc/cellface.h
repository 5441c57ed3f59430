// Cell + face operators on the lane-per-quadrature-column layout: the
// linear operators B and C of the Schur complement and the device
// functions every k_cellface-style kernel is built from.
//
// Each cell evaluates the flux of each of its four faces with itself as the
// "+" side and its outward normal. All the reference's face fluxes are
// antisymmetric under exchanging the sides (n -> -n), so this equals the
// reference's face-loop contribution to the cell without any atomics; every
// face flux is computed twice (once per side).
#pragma once
#include <type_traits>

#include "common.h"
#include "wave.h"


namespace atmo {

template <int NF>
struct FieldPtrs {
  const double* p[NF];
};

// Bottom/top faces of the structured mesh: neighbour face rows of this many
// fields are loaded together (one memory latency per group; larger groups
// cost the operators of many fields occupancy).
constexpr int FACE_ROW_GROUP = 4;

// Number of leading input fields a physics' face flux reads: Ph::NF_FACE
// where the physics declares it, all NF fields otherwise.
template <class Ph, class = void>
struct FaceFields {
  static constexpr int value = Ph::NF;
};
template <class Ph>
struct FaceFields<Ph, std::void_t<decltype(Ph::NF_FACE)>> {
  static constexpr int value = Ph::NF_FACE;
};

// B: euler_operator.h:1510-1617 (cell: submit_divergence, face: {p}[v.n],
// wall: p_D = p)
struct PhysB {
  static constexpr int NF = 1, NO = 2;
  static constexpr bool BDRY = true;
  // sparsity of the cell integrand: value term, and which outputs have an x / z gradient term
  static constexpr bool HAS_V = false;
  static constexpr unsigned GX = 0b01, GZ = 0b10;
  double cf;  // coeff*dt/Ma^2
  __device__ void cell(const double (&v)[NF], double (&V)[NO], double (&Gx)[NO], double (&Gz)[NO]) const {
    V[0] = V[1] = 0;
    Gx[0] = -cf * v[0];
    Gz[0] = 0;
    Gx[1] = 0;
    Gz[1] = -cf * v[0];
  }
  __device__ void face(const double (&a)[NF], const double (&b)[NF], double n0, double n1, double (&F)[NO]) const {
    const double avg = 0.5 * (a[0] + b[0]);
    F[0] = cf * avg * n0;
    F[1] = cf * avg * n1;
  }
  __device__ void bdry(const double (&a)[NF], double n0, double n1, double (&F)[NO]) const {
    const double pD = a[0];
    F[0] = cf * 0.5 * (a[0] + pD) * n0;
    F[1] = cf * 0.5 * (a[0] + pD) * n1;
  }
};

// C: euler_operator.h:1411-1507 (nonlinear in u through lambda)
struct PhysC {
  static constexpr int NF = 3, NO = 1;  // pfix, ux, uz
  static constexpr bool BDRY = false;
  static constexpr bool HAS_V = false;
  static constexpr unsigned GX = 0b1, GZ = 0b1;
  double cdt;
  // gamma factors, formed on the host in the order of the reference's
  // expressions (bit-identical to forming them per point)
  double cs;  // cdt*gamma/(gamma-1)
  double hg;  // 0.5*gamma/(gamma-1)
  double ig;  // 1/(gamma-1)
  void set(double cdt_, double gm) {
    cdt = cdt_;
    cs = cdt * gm / (gm - 1.0);
    hg = 0.5 * gm / (gm - 1.0);
    ig = 1.0 / (gm - 1.0);
  }
  __device__ void cell(const double (&v)[NF], double (&V)[NO], double (&Gx)[NO], double (&Gz)[NO]) const {
    const double s = -cs * v[0];
    V[0] = 0;
    Gx[0] = s * v[1];
    Gz[0] = s * v[2];
  }
  __device__ void face(const double (&a)[NF], const double (&b)[NF], double n0, double n1, double (&F)[NO]) const {
    const double ax = hg * (a[0] * a[1] + b[0] * b[1]);
    const double az = hg * (a[0] * a[2] + b[0] * b[2]);
    const double lam = fmax(fabs(a[1] * n0 + a[2] * n1), fabs(b[1] * n0 + b[2] * n1));
    const double jump = ig * (a[0] - b[0]);
    F[0] = cdt * (ax * n0 + az * n1 + 0.5 * lam * jump);
  }
  __device__ void bdry(const double (&)[NF], double, double, double (&F)[NO]) const { F[0] = 0; }
};

// D p - C u in one pass (the Schur complement without its inner solve,
// euler_operator.h:2374-2408): fields (pfix, ux, uz, p); D = M/(gamma-1)
// (:1384-1408) enters as a cell value term, C with the opposite sign.
struct PhysDminusC {
  static constexpr int NF = 4, NO = 1;
  static constexpr int NF_FACE = 3;  // the face flux is C's: p enters the cell term only
  static constexpr bool BDRY = false;
  static constexpr bool HAS_V = true;
  static constexpr unsigned GX = 0b1, GZ = 0b1;
  PhysC c;
  double cD;  // 1/(gamma-1)
  __device__ void cell(const double (&v)[NF], double (&V)[NO], double (&Gx)[NO], double (&Gz)[NO]) const {
    const double s = c.cs * v[0];
    V[0] = cD * v[3];
    Gx[0] = s * v[1];
    Gz[0] = s * v[2];
  }
  __device__ void face(const double (&a)[NF], const double (&b)[NF], double n0, double n1, double (&F)[NO]) const {
    const double ac[3] = {a[0], a[1], a[2]}, bc[3] = {b[0], b[1], b[2]};
    double Fc[1];
    c.face(ac, bc, n0, n1, Fc);
    F[0] = -Fc[0];
  }
  __device__ void bdry(const double (&)[NF], double, double, double (&F)[NO]) const { F[0] = 0; }
};

// LDS of one 64-lane block of a cell+face kernel. AMR (locally refined
// mesh): the fluxes of all four faces, per half of a split face, go through
// LDS with the face type of each cell; otherwise the bottom/top fluxes are
// folded into the column partials and only left/right go through LDS.
// Column partials: plane o holds output o's value/eta partials; an output
// with an x-gradient (bit o of Ph::GX) also has an xi plane, NO + (its rank
// among those outputs), so outputs without one take no LDS for it.
template <int K, class Ph, bool AMR>
struct CFSmem {
  static constexpr int ND = K + 1, NQ = 2 * K + 1, ND2 = ND * ND, CPW = 64 / NQ;
  static constexpr int NFACE = AMR ? 4 : 2, NHALF = AMR ? 2 : 1;
  static constexpr int CS = lds_cell_stride<K>();
  static constexpr int popc(unsigned x) { return x ? (int)(x & 1u) + popc(x >> 1) : 0; }
  static constexpr int NPLANE = Ph::NO + popc(Ph::GX & ((1u << Ph::NO) - 1u));
  // plane of output o's value/eta (d = 0) or xi (d = 1) partials
  static constexpr int plane(int o, int d) { return d == 0 ? o : Ph::NO + popc(Ph::GX & ((1u << o) - 1u)); }
  // operators on the structured mesh also stage the cells before and after the block's (their left/right
  // neighbour traces then come from LDS)
  static constexpr int HALO = !AMR ? 1 : 0;
  // the staged inputs are dead once the column phase has read them; the partials written at its end reuse
  // their space (after a barrier that follows the last read)
  union {
    double in[Ph::NF][(CPW + 2 * HALO) * CS];  // staged inputs, block cell w at (w + HALO)*CS
    double T[NPLANE][CPW][ND][NQ];             // column partials of the value/eta and xi test functions
  };
  double F[Ph::NO][NFACE][NHALF][CPW][NQ];   // face fluxes times the face weight
  int ft[AMR ? CPW : 1][4];                  // AMR: 0 no flux, 1 full face, 2 split face
};

// Trace of field q of a cell along face f at lane point c: Tab[i] weights
// along the face (the cell's rows for bottom/top, columns for left/right).
template <int K>
__device__ inline double face_trace(const double* __restrict__ u, int f, const double (&Tab)[K + 1]) {
  constexpr int ND = K + 1;
  double s = 0;
  if (f < 2) {
    const int col = f == 0 ? 0 : K;
#pragma unroll
    for (int j = 0; j < ND; ++j) s += Tab[j] * u[j * ND + col];
  } else {
    const int row = f == 2 ? 0 : K;
#pragma unroll
    for (int i = 0; i < ND; ++i) s += Tab[i] * u[row * ND + i];
  }
  return s;
}

// Column phase of a cell+face operator for the block's CPW cells: stages
// the input fields in LDS, evaluates cell and face terms per quadrature
// column (neighbour traces are read from memory; they are L2/MALL hits)
// and leaves the column partials in sh.T / sh.F. Ends with a block
// barrier.
struct NoPrefetch {
  __device__ void operator()() const {}
};
// post(): runs once the staged inputs are visible to the whole block, before the column phase (a caller
// that needs staged values after the column partials have taken over their LDS copies them out)
struct NoPost {
  __device__ void operator()() const {}
};

// pre(): issued after the staging loads and before the first barrier (a
// caller's own global loads, in flight during the column phase).
template <int K, class Ph, bool AMR, class Pre = NoPrefetch, class Post = NoPost>
__device__ inline void cellface_columns(const MeshDev& m, const Ph& ph, const FieldPtrs<Ph::NF>& in,
                                        CFSmem<K, Ph, AMR>& sh, const Pre& pre = Pre(), const Post& post = Post()) {
  using C = Cfg<K>;
  constexpr int ND = C::ND, NQ = C::NQ, ND2 = C::ND2;
  constexpr int NF = Ph::NF, NO = Ph::NO;
  LaneMap<K> L(m.ncell);
  // the cell's vertices and the lane's table rows are loaded before the staged fields, so the three share
  // one memory latency (an inactive lane reads cell 0's)
  const int vcell = L.active ? L.cell : 0;
  const VertsRaw vr = cell_verts_load(m, vcell);
  double Sc[ND], Dc[ND];
  lane_rows<K, double>(L.c, Sc, Dc);
  constexpr int HALO = CFSmem<K, Ph, AMR>::HALO, CS = CFSmem<K, Ph, AMR>::CS;
  stage_fields<K, NF, double, double, CS, HALO>(&sh.in[0][0], in.p, m.ncell);
  pre();
  __syncthreads();
  post();
  const CellVerts v = cell_verts_form(m, vcell, vr);
  double Tv[Ph::NO][ND], Td[Ph::NO][ND];
#pragma unroll
  for (int o = 0; o < NO; ++o)
#pragma unroll
    for (int j = 0; j < ND; ++j) Tv[o][j] = Td[o][j] = 0;
  if (L.active) {
    const ColGeo<double> g = col_geo<K, double>(v, L.c);
    const int off = (L.cw + HALO) * CS;
    double tv[NF][ND];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int j = 0; j < ND; ++j) {
        double s = 0;
#pragma unroll
        for (int i = 0; i < ND; ++i) s += Sc[i] * sh.in[f][off + j * ND + i];
        tv[f][j] = s;
      }
    const double wdx = g.wc * g.dx, wze = g.wc * g.ze, jc = wdx * g.ze;
    // Rolled point loop, table rows read from constant memory (scalar loads
    // per row): unrolled, every FP64 table entry becomes an SGPR-pair
    // literal and the kernels spill SGPRs into VGPR lanes.
#pragma unroll 1
    for (int r = 0; r < NQ; ++r) {
      const double* Sr = cT<K>.S[r];
      const double* Dr = cT<K>.D[r];
      double val[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        double s = 0;
#pragma unroll
        for (int j = 0; j < ND; ++j) s += Sr[j] * tv[f][j];
        val[f] = s;
      }
      double V[NO], Gx[NO], Gz[NO];
      ph.cell(val, V, Gx, Gz);
      const double wr = cT<K>.gw[r];
      const double zs = g.b0 + cT<K>.gq[r] * g.b1;
#pragma unroll
      for (int o = 0; o < NO; ++o) {
        const bool gx = (Ph::GX >> o) & 1, gz = (Ph::GZ >> o) & 1;
        // reference-space integrands times the weights: G_xi = ze*Gx, G_eta = dx*Gz - z_xi*Gx
        double Geta = 0.0;
        if (gx) Geta = -(g.wc * zs) * Gx[o];
        if (gz) Geta += wdx * Gz[o];
        Geta *= wr;
        const double Gxi = gx ? wr * wze * Gx[o] : 0.0;
        const double Vq = Ph::HAS_V ? V[o] * (jc * wr) : 0.0;
#pragma unroll
        for (int j = 0; j < ND; ++j) {
          if (Ph::HAS_V) Tv[o][j] += Sr[j] * Vq;
          if (gx || gz) Tv[o][j] += Dr[j] * Geta;
          if (gx) Td[o][j] += Sr[j] * Gxi;
        }
      }
    }
    if constexpr (!AMR && NF == 1) {
      const int cj = L.cell / m.nx, ci = L.cell - cj * m.nx;
      // neighbours of the faces f = 0 left i-1, 1 right i+1, 2 bottom j-1, 3 top j+1; a wall side reads its
      // own cell instead (the value is not used), so the loads need no branch
      bool inner[4], staged[2];
      long nbase[4];
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const int nj = side ? cj + 1 : cj - 1;
        inner[2 + side] = nj >= 0 && nj < m.nz;
        nbase[2 + side] = (long)((inner[2 + side] ? nj : cj) * m.nx + ci) * ND2;
        int ni = side ? ci + 1 : ci - 1;
        bool in_x = true;
        staged[side] = true;  // the cell before or after this one, staged next to it
        if (ni < 0 || ni >= m.nx) {
          staged[side] = false;
          if (m.periodic) ni = (ni + m.nx) % m.nx;
          else in_x = false;
        }
        inner[side] = in_x;
        nbase[side] = (long)(cj * m.nx + (in_x ? ni : ci)) * ND2;
      }
      // one field: the neighbour loads of all four faces are issued together (one memory latency for the
      // face phase); operators of more fields take the faces one at a time below, as four faces' loads
      // would not fit the registers their cell phase leaves
      constexpr int FG = 4;
#pragma unroll
      for (int g0 = 0; g0 < 4; g0 += FG) {
        double nb[FG][NF];
#pragma unroll
        for (int i = 0; i < FG; ++i) {
          const int f = (g0 + i + 2) & 3;
          if (f < 2 && staged[f & 1]) {
#pragma unroll
            for (int q = 0; q < NF; ++q) nb[i][q] = face_trace<K>(&sh.in[q][off + (f ? CS : -CS)], f ^ 1, Sc);
          } else {
#pragma unroll
            for (int q = 0; q < NF; ++q) nb[i][q] = face_trace<K>(in.p[q] + nbase[f], f ^ 1, Sc);
          }
        }
#pragma unroll
        for (int i = 0; i < FG; ++i) {
          const int f = (g0 + i + 2) & 3;
          const FaceGeo fg = face_geo(v, f);
          const double wl = fg.len * g.wc;
          double own[NF], F[NO];
          // own trace from LDS (this lane is face point c of the bottom/top faces, s = c of the left/right
          // ones): keeps tv dead after the cell integral
#pragma unroll
          for (int q = 0; q < NF; ++q) own[q] = face_trace<K>(&sh.in[q][off], f, Sc);
          if (inner[f]) {
            ph.face(own, nb[i], fg.n0, fg.n1, F);
          } else if (Ph::BDRY) {
            ph.bdry(own, fg.n0, fg.n1, F);
          } else {
#pragma unroll
            for (int o = 0; o < NO; ++o) F[o] = 0;
          }
          if (f >= 2) {
#pragma unroll
            for (int o = 0; o < NO; ++o) Tv[o][f == 3 ? K : 0] += F[o] * wl;
          } else {
#pragma unroll
            for (int o = 0; o < NO; ++o) sh.F[o][f][0][L.cw][L.c] = F[o] * wl;
          }
        }
      }
    } else if constexpr (!AMR) {
      const int cj = L.cell / m.nx, ci = L.cell - cj * m.nx;
      // bottom (j-1) and top (j+1) faces: this lane is face point c. Per face, the loads of the neighbour's
      // face row of a group of fields the flux reads are issued before the first of them is used (one
      // memory latency per group)
      constexpr int NFF = FaceFields<Ph>::value, RG = FACE_ROW_GROUP;
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const int f = 2 + side;
        const FaceGeo fg = face_geo(v, f);
        const double wl = fg.len * g.wc;
        double own[NF], nb[NF], F[NO];
        // own trace from LDS (as for the other faces): keeps tv dead after the cell integral
#pragma unroll
        for (int q = 0; q < NF; ++q) own[q] = face_trace<K>(&sh.in[q][off], f, Sc);
        const int nj = side ? cj + 1 : cj - 1;
        if (nj >= 0 && nj < m.nz) {
          // row K of the cell below, row 0 of the cell above; groups of up to RG fields
          const long nbase = (long)(nj * m.nx + ci) * ND2 + (side ? 0 : K * ND);
#pragma unroll
          for (int q0 = 0; q0 < NF; q0 += RG) {
            double row[RG][ND];
#pragma unroll
            for (int q = 0; q < RG; ++q)
#pragma unroll
              for (int i = 0; i < ND; ++i) row[q][i] = q0 + q < NFF ? in.p[q0 + q < NFF ? q0 + q : 0][nbase + i] : 0.0;
            // the trace of the row (face_trace's order of terms); fields the flux does not read are 0
#pragma unroll
            for (int q = 0; q < RG; ++q) {
              if (q0 + q >= NF) continue;
              double t = 0;
              if (q0 + q < NFF)
#pragma unroll
                for (int i = 0; i < ND; ++i) t += Sc[i] * row[q][i];
              nb[q0 + q] = t;
            }
          }
          ph.face(own, nb, fg.n0, fg.n1, F);
        } else if (Ph::BDRY) {
          ph.bdry(own, fg.n0, fg.n1, F);
        } else {
#pragma unroll
          for (int o = 0; o < NO; ++o) F[o] = 0;
        }
#pragma unroll
        for (int o = 0; o < NO; ++o) Tv[o][side ? K : 0] += F[o] * wl;
      }
      // left (i-1) and right (i+1) faces: this lane is face point s = c
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const int f = side;
        const FaceGeo fg = face_geo(v, f);
        const double wl = fg.len * g.wc;
        double own[NF], nb[NF], F[NO];
#pragma unroll
        for (int q = 0; q < NF; ++q) own[q] = face_trace<K>(&sh.in[q][off], f, Sc);
        int ni = side ? ci + 1 : ci - 1;
        bool inner = true, staged = HALO == 1;
        if (ni < 0 || ni >= m.nx) {
          staged = false;
          if (m.periodic) ni = (ni + m.nx) % m.nx;
          else inner = false;
        }
        if (inner) {
          if (staged) {
            // the neighbour is the cell before or after this one, staged next to it
#pragma unroll
            for (int q = 0; q < NF; ++q) nb[q] = face_trace<K>(&sh.in[q][off + (side ? CS : -CS)], f ^ 1, Sc);
          } else {
            const long nbase = (long)(cj * m.nx + ni) * ND2;
#pragma unroll
            for (int q = 0; q < NF; ++q) nb[q] = face_trace<K>(in.p[q] + nbase, f ^ 1, Sc);
          }
          ph.face(own, nb, fg.n0, fg.n1, F);
        } else if (Ph::BDRY) {
          ph.bdry(own, fg.n0, fg.n1, F);
        } else {
#pragma unroll
          for (int o = 0; o < NO; ++o) F[o] = 0;
        }
#pragma unroll
        for (int o = 0; o < NO; ++o) sh.F[o][side][0][L.cw][L.c] = F[o] * wl;
      }
    } else {
      // locally refined mesh: per face, full (same-size or larger neighbour)
      // or split in two halves (two smaller neighbours)
      double Sh0[ND], Sh1[ND];
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        Sh0[i] = cT<K>.Sh[0][L.c][i];
        Sh1[i] = cT<K>.Sh[1][L.c][i];
      }
      const int* rec = m.cface + 16 * (long)L.cell;
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const int type = rec[4 * f], n0 = rec[4 * f + 1], n1 = rec[4 * f + 2], half = rec[4 * f + 3];
        const FaceGeo fg = face_geo(v, f);
        double own[NF], nb[NF], F[NO];
        if (type != FACE_FINER) {
          const double wl = fg.len * g.wc;
#pragma unroll
          for (int q = 0; q < NF; ++q) own[q] = face_trace<K>(&sh.in[q][off], f, Sc);
          if (type == FACE_BDRY) {
            if (Ph::BDRY) ph.bdry(own, fg.n0, fg.n1, F);
            else
#pragma unroll
              for (int o = 0; o < NO; ++o) F[o] = 0;
          } else {
            // same-size neighbour at point c, or the larger neighbour at (half + gq[c])/2 of its face
            const double* Tn = type == FACE_SAME ? Sc : (half ? Sh1 : Sh0);
            double Tab[ND];
#pragma unroll
            for (int i = 0; i < ND; ++i) Tab[i] = Tn[i];
#pragma unroll
            for (int q = 0; q < NF; ++q) nb[q] = face_trace<K>(in.p[q] + (long)n0 * ND2, f ^ 1, Tab);
            ph.face(own, nb, fg.n0, fg.n1, F);
          }
#pragma unroll
          for (int o = 0; o < NO; ++o) sh.F[o][f][0][L.cw][L.c] = F[o] * wl;
        } else {
          const double wl = 0.5 * fg.len * g.wc;
#pragma unroll
          for (int hh = 0; hh < 2; ++hh) {
            double Tab[ND];
#pragma unroll
            for (int i = 0; i < ND; ++i) Tab[i] = hh ? Sh1[i] : Sh0[i];
#pragma unroll
            for (int q = 0; q < NF; ++q) {
              own[q] = face_trace<K>(&sh.in[q][off], f, Tab);
              nb[q] = face_trace<K>(in.p[q] + (long)(hh ? n1 : n0) * ND2, f ^ 1, Sc);
            }
            ph.face(own, nb, fg.n0, fg.n1, F);
#pragma unroll
            for (int o = 0; o < NO; ++o) sh.F[o][f][hh][L.cw][L.c] = F[o] * wl;
          }
        }
        if (L.c == 0) sh.ft[L.cw][f] = type == FACE_FINER ? 2 : 1;
      }
    }
  }
  // the column partials reuse the staged inputs' LDS: every lane's last read of the inputs (own and
  // neighbour face traces) is ordered before the first write of a partial
  __syncthreads();
  if (L.active) {
#pragma unroll
    for (int o = 0; o < NO; ++o)
#pragma unroll
      for (int j = 0; j < ND; ++j) {
        sh.T[CFSmem<K, Ph, AMR>::plane(o, 0)][L.cw][j][L.c] = Tv[o][j];
        if ((Ph::GX >> o) & 1) sh.T[CFSmem<K, Ph, AMR>::plane(o, 1)][L.cw][j][L.c] = Td[o][j];
      }
  }
  __syncthreads();
}

// Table rows an output thread needs for block-local dof t (ii, jj = its x
// and z node). When ND2 divides 64 (k = 1, 3) they do not change across a
// thread's passes t, t+64, ..., so callers load them once.
template <int K>
struct OutRows {
  static constexpr int NQ = 2 * K + 1;
  double Si[NQ], Di[NQ], Sj[NQ];
};
template <int K>
__device__ inline OutRows<K> out_rows(int t) {
  constexpr int ND = K + 1, NQ = 2 * K + 1, ND2 = ND * ND;
  const int o2 = t % ND2, jj = o2 / ND, ii = o2 - jj * ND;
  OutRows<K> R;
#pragma unroll
  for (int c = 0; c < NQ; ++c) {
    R.Si[c] = cT<K>.St[ii][c];
    R.Di[c] = cT<K>.Dt[ii][c];
    R.Sj[c] = cT<K>.St[jj][c];
  }
  return R;
}
template <int K>
constexpr bool rows_invariant() {
  return 64 % ((K + 1) * (K + 1)) == 0;
}

// Output phase: component o of block-local dof t = cw*ND2 + jj*ND + ii.
template <int K, class Ph, bool AMR>
__device__ inline double cellface_out(const CFSmem<K, Ph, AMR>& sh, int o, int t, const OutRows<K>& R) {
  constexpr int ND = K + 1, NQ = 2 * K + 1, ND2 = ND * ND;
  const int cw = t / ND2, o2 = t - cw * ND2, jj = o2 / ND, ii = o2 - jj * ND;
  const bool gx = (Ph::GX >> o) & 1;
  double acc = 0;
#pragma unroll
  for (int c = 0; c < NQ; ++c) {
    acc += R.Si[c] * sh.T[CFSmem<K, Ph, AMR>::plane(o, 0)][cw][jj][c];
    if (gx) acc += R.Di[c] * sh.T[CFSmem<K, Ph, AMR>::plane(o, 1)][cw][jj][c];
  }
  if constexpr (!AMR) {
    if (ii == 0 || ii == K) {
      const int side = ii == 0 ? 0 : 1;
#pragma unroll
      for (int c = 0; c < NQ; ++c) acc += R.Sj[c] * sh.F[o][side][0][cw][c];
    }
  } else {
    // faces through this dof: left/right weigh along jj, bottom/top along ii
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const bool on = f == 0 ? ii == 0 : f == 1 ? ii == K : f == 2 ? jj == 0 : jj == K;
      if (!on) continue;
      const int along = f < 2 ? jj : ii;
      if (sh.ft[cw][f] == 2) {
#pragma unroll
        for (int hh = 0; hh < 2; ++hh)
#pragma unroll
          for (int c = 0; c < NQ; ++c) acc += cT<K>.ShT[hh][along][c] * sh.F[o][f][hh][cw][c];
      } else {
#pragma unroll
        for (int c = 0; c < NQ; ++c) acc += cT<K>.St[along][c] * sh.F[o][f][0][cw][c];
      }
    }
  }
  return acc;
}

}  // namespace atmo
