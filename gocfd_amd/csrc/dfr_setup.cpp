// Host-side setup: mesh ingest, topology, operators, geometry, device layout.
#include "dfr_setup.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>
#include <unordered_map>

#include "gocfd_euler2d.h"
#include "wsj_nodes.h"

namespace gocfd {

// ===========================================================================
// Small dense linear algebra (row-major)
// ===========================================================================
namespace {

struct Mat {
  int r = 0, c = 0;
  std::vector<double> a;
  Mat() = default;
  Mat(int r_, int c_) : r(r_), c(c_), a(size_t(r_) * c_, 0.0) {}
  double &operator()(int i, int j) { return a[size_t(i) * c + j]; }
  double operator()(int i, int j) const { return a[size_t(i) * c + j]; }
};

Mat matmul(const Mat &A, const Mat &B) {
  Mat C(A.r, B.c);
  for (int i = 0; i < A.r; ++i)
    for (int l = 0; l < A.c; ++l) {
      const double x = A(i, l);
      for (int j = 0; j < B.c; ++j) C(i, j) += x * B(l, j);
    }
  return C;
}

// Inverse through an LU factorisation with partial pivoting (the reference
// uses gonum Getrf/Getri, utils/matrix_extended.go:953-996).
Mat inverse(const Mat &A, const char *what) {
  const int n = A.r;
  Mat LU = A;
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(LU(i, k)) > std::fabs(LU(p, k))) p = i;
    if (std::fabs(LU(p, k)) < 1e-300)
      throw SetupError(GOCFD_ERR_SETUP, std::string("singular matrix: ") + what);
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(LU(k, j), LU(p, j));
    for (int i = k + 1; i < n; ++i) {
      LU(i, k) /= LU(k, k);
      for (int j = k + 1; j < n; ++j) LU(i, j) -= LU(i, k) * LU(k, j);
    }
  }
  Mat X(n, n);
  for (int col = 0; col < n; ++col) {
    std::vector<double> b(n, 0.0);
    b[col] = 1.0;
    for (int k = 0; k < n; ++k) std::swap(b[k], b[piv[k]]);
    for (int i = 1; i < n; ++i)
      for (int j = 0; j < i; ++j) b[i] -= LU(i, j) * b[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) b[i] -= LU(i, j) * b[j];
      b[i] /= LU(i, i);
    }
    for (int i = 0; i < n; ++i) X(i, col) = b[i];
  }
  return X;
}

// Normalised Jacobi polynomial P_n^(a,b) and its derivative (DG1D/elements.go:257-328).
double jacobi(double x, double a, double b, int n) {
  const double ab = a + b;
  const double g0 = std::tgamma(a + 1) * std::tgamma(b + 1) * std::pow(2.0, ab + 1) / (ab + 1) /
                    std::tgamma(ab + 1);
  double p0 = 1.0 / std::sqrt(g0);
  if (n == 0) return p0;
  const double g1 = (a + 1) * (b + 1) * g0 / (ab + 3);
  double p1 = ((ab + 2) * x / 2 + (a - b) / 2) / std::sqrt(g1);
  if (n == 1) return p1;
  double aold = 2.0 * std::sqrt((a + 1) * (b + 1) / (ab + 3)) / (ab + 2);
  for (int i = 1; i < n; ++i) {
    const double h1 = 2.0 * i + ab;
    const double anew = 2.0 / (h1 + 2) *
                        std::sqrt((i + 1.0) * (i + 1 + ab) * (i + 1 + a) * (i + 1 + b) / (h1 + 1) / (h1 + 3));
    const double bnew = -(a * a - b * b) / h1 / (h1 + 2);
    const double p2 = ((x - bnew) * p1 - aold * p0) / anew;
    p0 = p1;
    p1 = p2;
    aold = anew;
  }
  return p1;
}
double djacobi(double x, double a, double b, int n) {
  return n == 0 ? 0.0 : std::sqrt(n * (n + a + b + 1.0)) * jacobi(x, a + 1, b + 1, n - 1);
}

// Orthonormal PKD basis term (i,j) on the reference triangle and its
// gradient (DG2D/basis_polynomials.go:72-153).
void pkd(double r, double s, int i, int j, double *val, double *dr, double *ds) {
  const double a = (s != 1.0) ? 2 * (1 + r) / (1 - s) - 1 : -1.0;
  const double b = s;
  const double fa = jacobi(a, 0, 0, i), gb = jacobi(b, 2.0 * i + 1, 0, j);
  const double w = 0.5 * (1 - b);
  if (val) val[0] = std::sqrt(2.0) * fa * gb * std::pow(1 - b, i);
  if (dr || ds) {
    const double dfa = djacobi(a, 0, 0, i), dgb = djacobi(b, 2.0 * i + 1, 0, j);
    const double scale = std::pow(2.0, i + 0.5);
    const double wim1 = (i > 0) ? std::pow(w, i - 1) : 1.0;
    if (dr) dr[0] = dfa * gb * wim1 * scale;
    if (ds) {
      double t = 0.5 * dfa * gb * (1 + a) * wim1;
      double u = dgb * std::pow(w, i);
      if (i > 0) u -= 0.5 * i * gb * wim1;
      ds[0] = (t + fa * u) * scale;
    }
  }
}

int np2d(int n) { return (n + 1) * (n + 2) / 2; }

Mat vandermonde(int n, const std::vector<double> &R, const std::vector<double> &S, size_t off,
                int npts) {
  Mat V(npts, np2d(n));
  for (int p = 0; p < npts; ++p) {
    int col = 0;
    for (int i = 0; i <= n; ++i)
      for (int j = 0; j <= n - i; ++j) pkd(R[off + p], S[off + p], i, j, &V(p, col++), nullptr, nullptr);
  }
  return V;
}

// Lagrange interpolant set on a 2D node set: value/gradient of l_j at (r,s).
struct Lagrange2D {
  int n;
  Mat Vinv;
  void eval(double r, double s, int j, double &v, double &gr, double &gs) const {
    v = gr = gs = 0;
    int m = 0;
    for (int i = 0; i <= n; ++i)
      for (int k = 0; k <= n - i; ++k, ++m) {
        double pv, pr, ps;
        pkd(r, s, i, k, &pv, &pr, &ps);
        v += pv * Vinv(m, j);
        gr += pr * Vinv(m, j);
        gs += ps * Vinv(m, j);
      }
  }
};

struct Lagrange1D {
  int n;
  Mat Vinv;
  void eval(double t, int j, double &v, double &d) const {
    v = d = 0;
    for (int i = 0; i <= n; ++i) {
      v += jacobi(t, 0, 0, i) * Vinv(i, j);
      d += djacobi(t, 0, 0, i) * Vinv(i, j);
    }
  }
};

// Edge point parameters of the RT element (GetOptimizedEdgePoints,
// DG2D/raviart_thomas_element.go:377-400), indexed by RT order.
const std::vector<double> &rt_edge_params(int P) {
  static const std::vector<std::vector<double>> t = {
      {},
      {-0.38490018, 0.38490018},
      {-0.028, 0.000, 0.028},
      {-0.482, -0.161, 0.161, 0.482},
      {-0.586, -0.293, 0.000, 0.293, 0.586},
      {-0.634, -0.381, -0.127, 0.127, 0.381, 0.634},
      {-0.671, -0.452, -0.219, 0.000, 0.219, 0.452, 0.671},
      {-0.808, -0.578, -0.343, -0.113, 0.113, 0.343, 0.578, 0.808},
      {-0.692, -0.523, -0.350, -0.176, 0.000, 0.176, 0.350, 0.523, 0.692}};
  if (P < 1 || P > 8) throw SetupError(GOCFD_ERR_SETUP, "no RT edge points for this order");
  return t[P];
}

}  // namespace

// ===========================================================================
// Operators: NewRTElement(N+1, SimplexRTBasis) + NewLagrangeElement2D(N)
// (raviart_thomas_element.go:192-297, rt_basis_simplex.go:17-273,
//  basis_polynomial_construction.go:331-359, lagrange_element.go:54-92)
// ===========================================================================
Operators build_operators(int N) {
  if (N < 0 || N > 7) throw SetupError(GOCFD_ERR_ARG, "Polynomial order must be 0..7");
  Operators o;
  const int P = N + 1;
  o.N = N;
  o.NpInt = np2d(N);
  o.NpEdge = P + 1;
  o.NpFlux = (P + 1) * (P + 3);
  const int NpInt = o.NpInt, NpEdge = o.NpEdge, NpFlux = o.NpFlux;

  // Solution nodes: WSJ barycentrics -> (r,s) (nodes2d_williams_shunn_jameson.go:156-163)
  const double(*w)[4] = wsj::table(N);
  o.solR.resize(NpInt); o.solS.resize(NpInt); o.solW.resize(NpInt);
  for (int i = 0; i < NpInt; ++i) {
    o.solR[i] = -w[i][0] + w[i][1] - w[i][2];
    o.solS[i] = -w[i][0] - w[i][1] + w[i][2];
    o.solW[i] = w[i][3];
  }
  // RT points: interior nodes twice, then bottom / hypotenuse / left edges (ExtendGeomToRT :429-496)
  o.R.resize(NpFlux); o.S.resize(NpFlux);
  for (int i = 0; i < NpInt; ++i) {
    o.R[i] = o.R[NpInt + i] = o.solR[i];
    o.S[i] = o.S[NpInt + i] = o.solS[i];
  }
  const auto &g = rt_edge_params(P);
  for (int i = 0; i < NpEdge; ++i) {
    const double t = 0.5 * (g[i] + 1);
    o.R[2 * NpInt + i] = g[i];
    o.S[2 * NpInt + i] = -1;
    o.R[2 * NpInt + NpEdge + i] = 1 - 2 * t;
    o.S[2 * NpInt + NpEdge + i] = -1 + 2 * t;
    o.R[2 * NpInt + 2 * NpEdge + i] = -1;
    o.S[2 * NpInt + 2 * NpEdge + i] = -g[i];
  }
  // Interior polynomial multipliers: Lagrange basis of order N on the solution nodes
  Lagrange2D L2{N, inverse(vandermonde(N, o.solR, o.solS, 0, NpInt), "interior Vandermonde")};
  // Edge polynomial multipliers: 1D Lagrange basis of order P on the edge parameters
  Mat V1(NpEdge, NpEdge);
  for (int i = 0; i < NpEdge; ++i)
    for (int j = 0; j < NpEdge; ++j) V1(i, j) = jacobi(g[i], 0, 0, j);
  Lagrange1D L1{P, inverse(V1, "edge Vandermonde")};

  // RT basis function j: value (vector) and divergence at (r,s)
  auto phi = [&](int j, double r, double s, double v[2], double &div) {
    const double rp = 0.5 * (r + 1), rm = 0.5 * (r - 1), sp = 0.5 * (s + 1), sm = 0.5 * (s - 1);
    double base[2], bdiv, pv, gr = 0, gs = 0;
    if (j < 2 * NpInt) {
      const bool e4 = j < NpInt;
      if (e4) { base[0] = sp * rp; base[1] = sp * sm; bdiv = (3 * s + 1) / 4; }
      else { base[0] = rp * rm; base[1] = rp * sp; bdiv = (3 * r + 1) / 4; }
      L2.eval(r, s, e4 ? j : j - NpInt, pv, gr, gs);
    } else {
      const int jj = j - 2 * NpInt, edge = jj / NpEdge, m = jj % NpEdge;
      double d;
      if (edge == 0) {        // bottom: parameter r
        L1.eval(r, m, pv, d); gr = d;
        base[0] = rp; base[1] = sm; bdiv = 1;
      } else if (edge == 1) { // hypotenuse: parameter s
        L1.eval(s, m, pv, d); gs = d;
        base[0] = std::sqrt(2.0) * rp; base[1] = std::sqrt(2.0) * sp; bdiv = std::sqrt(2.0);
      } else {                // left: parameter -s
        L1.eval(-s, m, pv, d); gs = -d;
        base[0] = rm; base[1] = sp; bdiv = 1;
      }
    }
    v[0] = base[0] * pv;
    v[1] = base[1] * pv;
    div = pv * bdiv + base[0] * gr + base[1] * gs;
  };
  // DOF unit vectors (raviart_thomas_element.go:224-234)
  auto dofvec = [&](int i, double b[2]) {
    const double h = 0.5 * std::sqrt(2.0);
    if (i < NpInt) { b[0] = 1; b[1] = 0; }
    else if (i < 2 * NpInt) { b[0] = 0; b[1] = 1; }
    else if (i < 2 * NpInt + NpEdge) { b[0] = 0; b[1] = -1; }
    else if (i < 2 * NpInt + 2 * NpEdge) { b[0] = h; b[1] = h; }
    else { b[0] = -1; b[1] = 0; }
  };
  Mat V(NpFlux, NpFlux), DB(NpFlux, NpFlux);
  for (int i = 0; i < NpFlux; ++i) {
    double b[2];
    dofvec(i, b);
    for (int j = 0; j < NpFlux; ++j) {
      double v[2], d;
      phi(j, o.R[i], o.S[i], v, d);
      V(i, j) = b[0] * v[0] + b[1] * v[1];
      DB(i, j) = d;
    }
  }
  Mat Div = matmul(DB, inverse(V, "RT Vandermonde"));
  o.Div = Div.a;
  o.DivInt.assign(Div.a.begin(), Div.a.begin() + size_t(NpInt) * NpFlux);

  // Solution element: V, Vinv, Mass = V^T diag(W) V
  Mat Vs = vandermonde(N, o.solR, o.solS, 0, NpInt);
  Mat Vsi = inverse(Vs, "solution Vandermonde");
  o.V = Vs.a;
  o.Vinv = Vsi.a;
  o.Mass.assign(size_t(NpInt) * NpInt, 0.0);
  for (int i = 0; i < NpInt; ++i)
    for (int j = 0; j < NpInt; ++j) {
      double acc = 0;
      for (int l = 0; l < NpInt; ++l) acc += Vs(l, i) * o.solW[l] * Vs(l, j);
      o.Mass[size_t(i) * NpInt + j] = acc;
    }
  // FluxEdgeInterp: solution basis evaluated at the 3*NpEdge edge points (basis_polynomials.go:49-53)
  Mat Ve = vandermonde(N, o.R, o.S, size_t(2) * NpInt, 3 * NpEdge);
  o.Interp = matmul(Ve, Vsi).a;
  return o;
}

std::vector<double> interp_matrix(const Operators &o, int npts, const double *r, const double *s) {
  const std::vector<double> R(r, r + npts), S(s, s + npts);
  Mat Vp = vandermonde(o.N, R, S, 0, npts);
  Mat Vi(o.NpInt, o.NpInt);
  Vi.a = o.Vinv;
  return matmul(Vp, Vi).a;
}

// ===========================================================================
// Mesh ingest
// ===========================================================================
namespace {
std::string lower_trim(std::string s) {
  auto b = s.find_first_not_of(" \t\r\n");
  auto e = s.find_last_not_of(" \t\r\n");
  s = (b == std::string::npos) ? std::string() : s.substr(b, e - b + 1);
  for (auto &ch : s) ch = char(std::tolower((unsigned char)ch));
  return s;
}

struct LineReader {
  std::ifstream f;
  std::string path;
  explicit LineReader(const std::string &p) : f(p), path(p) {
    if (!f) throw SetupError(GOCFD_ERR_MESH, "unable to open file " + p);
  }
  std::string line() {
    std::string l;
    if (!std::getline(f, l)) throw SetupError(GOCFD_ERR_MESH, "early end of file: " + path);
    if (!l.empty() && l.back() == '\r') l.pop_back();
    return l;
  }
  void skip(int n) { for (int i = 0; i < n; ++i) line(); }
  // SU2 "KEY= value" token, skipping '%' comment lines (readSU2Grid.go:104-154)
  std::string token() {
    for (;;) {
      std::string l = line();
      auto p = l.find_first_not_of(' ');
      if (p != std::string::npos && l[p] == '%') continue;
      auto eq = l.find('=');
      if (eq == std::string::npos)
        throw SetupError(GOCFD_ERR_MESH, "badly formed input line [" + l + "], should have an =");
      return l.substr(eq + 1);
    }
  }
  int number() {
    std::istringstream ss(token());
    int v;
    if (!(ss >> v)) throw SetupError(GOCFD_ERR_MESH, "unable to read number from token");
    return v;
  }
};

// ReadSU2 (readfiles/readSU2Grid.go:156-177): NDIME, NELEM (type 5), NPOIN, NMARK (type 3)
Mesh read_su2(const std::string &path) {
  LineReader rd(path);
  Mesh m;
  (void)rd.number();
  m.K = rd.number();
  m.EToV.resize(size_t(m.K) * 3);
  for (int k = 0; k < m.K; ++k) {
    std::istringstream ss(rd.line());
    int t, a, b, c;
    if (!(ss >> t >> a >> b >> c)) throw SetupError(GOCFD_ERR_MESH, "unable to read vertices");
    if (t != 5) throw SetupError(GOCFD_ERR_MESH, "unable to deal with non-triangular elements right now");
    m.EToV[3 * k] = a; m.EToV[3 * k + 1] = b; m.EToV[3 * k + 2] = c;
  }
  m.Nv = rd.number();
  m.VX.resize(m.Nv); m.VY.resize(m.Nv);
  for (int i = 0; i < m.Nv; ++i) {
    std::istringstream ss(rd.line());
    if (!(ss >> m.VX[i] >> m.VY[i])) throw SetupError(GOCFD_ERR_MESH, "unable to read coordinates");
  }
  int nmark = 0;
  try { nmark = rd.number(); } catch (const SetupError &) { nmark = 0; }
  for (int n = 0; n < nmark; ++n) {
    std::istringstream ls(rd.token());
    std::string label;
    ls >> label;
    const int ne = rd.number();
    std::vector<std::pair<int, int>> e(ne);
    for (int i = 0; i < ne; ++i) {
      std::istringstream ss(rd.line());
      int t, a, b;
      if (!(ss >> t >> a >> b)) throw SetupError(GOCFD_ERR_MESH, "unable to read boundary edge");
      if (t != 3) throw SetupError(GOCFD_ERR_MESH, "BCs should only contain line elements in 2D");
      e[i] = {a, b};
    }
    m.add_tag(lower_trim(label), std::move(e));
  }
  return m;
}

// ReadGambit2d (readfiles/readGambitGrid.go:21-165)
Mesh read_gambit(const std::string &path) {
  LineReader rd(path);
  Mesh m;
  rd.skip(6);
  int nmats, nbcs, nsd, dum;
  {
    std::istringstream ss(rd.line());
    if (!(ss >> m.Nv >> m.K >> nmats >> nbcs >> nsd >> dum))
      throw SetupError(GOCFD_ERR_MESH, "unable to read Gambit header");
  }
  if (nsd != 2) throw SetupError(GOCFD_ERR_MESH, "only 2D Gambit meshes are supported");
  rd.skip(2);
  m.VX.assign(m.Nv, 0.0); m.VY.assign(m.Nv, 0.0);
  for (int i = 0; i < m.Nv; ++i) {
    std::istringstream ss(rd.line());
    int ind; double x, y;
    if (!(ss >> ind >> x >> y) || ind < 1 || ind > m.Nv) throw SetupError(GOCFD_ERR_MESH, "bad vertex line");
    m.VX[ind - 1] = x; m.VY[ind - 1] = y;
  }
  rd.skip(2);
  m.EToV.assign(size_t(m.K) * 3, 0);
  for (int i = 0; i < m.K; ++i) {
    std::istringstream ss(rd.line());
    int ind, typ, nf, a, b, c;
    if (!(ss >> ind >> typ >> nf >> a >> b >> c) || ind < 1 || ind > m.K)
      throw SetupError(GOCFD_ERR_MESH, "bad element line");
    m.EToV[3 * (ind - 1)] = a - 1; m.EToV[3 * (ind - 1) + 1] = b - 1; m.EToV[3 * (ind - 1) + 2] = c - 1;
  }
  rd.skip(2);
  for (int gi = 0; gi < nmats; ++gi) {
    std::string h = rd.line();
    int gn = 0, elnum = 0;
    double mv = 0;
    if (std::sscanf(h.c_str(), "GROUP: %d ELEMENTS: %d MATERIAL: %lf", &gn, &elnum, &mv) != 3)
      throw SetupError(GOCFD_ERR_MESH, "bad material group header");
    rd.skip(2);
    rd.skip(elnum / 10 + (elnum % 10 ? 1 : 0));
    rd.skip(2);
  }
  for (int i = 0; i < nbcs; ++i) {
    if (i != 0) rd.skip(1);
    std::istringstream ss(rd.line());
    std::string name;
    ss >> name;
    name = lower_trim(name);
    int nfaces = 0;
    if (name == "cyl") { double pf; ss >> pf >> nfaces; }
    else { int id; ss >> id >> nfaces; }
    std::vector<std::pair<int, int>> e(nfaces);
    for (int j = 0; j < nfaces; ++j) {
      std::istringstream fs(rd.line());
      int kp1, n2, fp1;
      if (!(fs >> kp1 >> n2 >> fp1) || kp1 < 1 || kp1 > m.K) throw SetupError(GOCFD_ERR_MESH, "bad BC face line");
      const int *v = &m.EToV[3 * (kp1 - 1)];
      if (fp1 == 1) e[j] = {v[0], v[1]};
      else if (fp1 == 2) e[j] = {v[1], v[2]};
      else e[j] = {v[2], v[0]};
    }
    m.add_tag(name, std::move(e));
    rd.skip(1);
  }
  return m;
}
}  // namespace

void Mesh::add_tag(const std::string &tag, std::vector<std::pair<int, int>> e) {
  for (auto &t : tags)
    if (t.tag == tag) {  // BCMAP.AddEdges appends repeated tags (types/cfd.go:86-104)
      t.edges.insert(t.edges.end(), e.begin(), e.end());
      return;
    }
  tags.push_back({tag, std::move(e)});
}

Mesh read_mesh(const std::string &path) {
  std::string p = path;
  while (!p.empty() && p.back() == ' ') p.pop_back();
  if (p.size() < 4 || p[p.size() - 4] != '.')
    throw SetupError(GOCFD_ERR_ARG, "unable to determine file type from name: " + path);
  const std::string ext = p.substr(p.size() - 3);
  if (ext == "su2") return read_su2(p);
  if (ext == "neu") return read_gambit(p);
  throw SetupError(GOCFD_ERR_ARG, "unsupported file type: " + path);
}

int bc_flag_of_tag(const std::string &tag) {
  std::string base = tag;
  auto d = base.find('-');
  if (d != std::string::npos && d > 0) base = base.substr(0, d);
  static const std::map<std::string, int> names = {
      {"inflow", BC_In}, {"in", BC_In}, {"out", BC_Out}, {"outflow", BC_Out}, {"wall", BC_Wall},
      {"far", BC_Far}, {"cyl", BC_Cyl}, {"dirichlet", BC_Dirichlet}, {"neuman", BC_Neuman},
      {"slip", BC_Slip}, {"periodic", BC_Periodic}, {"periodicreversed", BC_PeriodicReversed}};
  auto it = names.find(base);
  return it == names.end() ? -1 : it->second;
}

// ===========================================================================
// Topology (DG2D/triangulation.go:19-171, types/elemental.go:14-241)
// ===========================================================================
namespace {
uint64_t key_of(int a, int b) {
  const uint64_t lo = uint64_t(std::min(a, b)), hi = uint64_t(std::max(a, b));
  return lo + (hi << 32);
}

// Outward-normal orientation of an edge seen from a triangle that traverses
// it with direction flag `rev` (GetEdgeCoordinates, triangulation.go:230-235):
// returns the start/end vertices whose difference gives n = (-dy, dx)/|d|.
void oriented(uint64_t key, int rev, int &a, int &b) {
  const int lo = int(key & 0xffffffffu), hi = int(key >> 32);
  a = rev ? lo : hi;
  b = rev ? hi : lo;
}

double iinii_of(const Discretisation &d, uint64_t key, int rev, int enr) {
  int a, b;
  oriented(key, rev, a, b);
  const double dx = d.VX[b] - d.VX[a], dy = d.VY[b] - d.VY[a];
  const double len = std::sqrt(dy * dy + dx * dx);
  return enr == 1 ? len / (2. * std::sqrt(2.0)) : len / 2.;
}

// Curve.ReOrder/AssembleCurve: order a list of segments into a connected
// chain starting from the first (or, reversed, the last) listed segment.
// A list that does not start at a chain end starts from the end vertex with
// the smaller id (the reference picks a map-order dependent end).
std::vector<std::pair<int, int>> reorder_chain(const std::vector<std::pair<int, int>> &c, bool reverse) {
  const int l = int(c.size());
  std::unordered_map<int, std::vector<int>> at;
  for (int i = 0; i < l; ++i) {
    at[c[i].first].push_back(i);
    at[c[i].second].push_back(i);
  }
  std::vector<int> ends;
  for (auto &kv : at) {
    if (kv.second.size() > 2) throw SetupError(GOCFD_ERR_MESH, "periodic curve branches");
    if (kv.second.size() == 1) ends.push_back(kv.first);
  }
  if (ends.size() != 2) throw SetupError(GOCFD_ERR_MESH, "unable to find two unconnected endKeys");
  std::sort(ends.begin(), ends.end());
  int start = reverse ? l - 1 : 0;
  if (at[ends[0]][0] != start && at[ends[1]][0] != start) start = at[ends[0]][0];
  std::vector<std::pair<int, int>> out{c[start]};
  int endv = at[c[start].first].size() == 2 ? c[start].first : c[start].second;
  int cur = start;
  while (at[endv].size() == 2 && int(out.size()) < l) {
    const int nxt = at[endv][0] != cur ? at[endv][0] : at[endv][1];
    out.push_back(c[nxt]);
    endv = (c[nxt].first == endv) ? c[nxt].second : c[nxt].first;
    cur = nxt;
  }
  if (int(out.size()) != l) throw SetupError(GOCFD_ERR_MESH, "periodic curve is not a single chain");
  return out;
}
}  // namespace

Discretisation build_discretisation(const Mesh &mesh, int N) {
  Discretisation d;
  d.N = N;
  d.K = mesh.K;
  d.Nv = mesh.Nv;
  d.ops = build_operators(N);
  d.VX = mesh.VX;
  d.VY = mesh.VY;
  d.EToV = mesh.EToV;
  const int K = d.K;
  for (int v : d.EToV)
    if (v < 0 || v >= d.Nv) throw SetupError(GOCFD_ERR_MESH, "element vertex index out of range");

  // Edges: conn 0 is the first (k, local edge) visit in element order.
  struct Visit { uint64_t key; int seq; };
  std::vector<Visit> visits(size_t(3) * K);
  for (int k = 0; k < K; ++k)
    for (int e = 0; e < 3; ++e)
      visits[3 * k + e] = {key_of(d.EToV[3 * k + e], d.EToV[3 * k + (e + 1) % 3]), 3 * k + e};
  std::sort(visits.begin(), visits.end(), [](const Visit &a, const Visit &b) {
    return a.key != b.key ? a.key < b.key : a.seq < b.seq;
  });
  d.EtoEdge.assign(size_t(3) * K, -1);
  for (const auto &v : visits) {
    if (d.edges.empty() || d.edges.back().key != v.key) {
      d.edges.emplace_back();
      d.edges.back().key = v.key;
    }
    Edge &E = d.edges.back();
    if (E.nconn > 1) throw SetupError(GOCFD_ERR_MESH, "incorrect edge construction, more than two connected triangles");
    const int k = v.seq / 3, e = v.seq % 3, c = E.nconn++;
    const int a = d.EToV[3 * k + e], b = d.EToV[3 * k + (e + 1) % 3];
    E.tri[c] = k; E.enr[c] = e; E.dir[c] = a > b ? 1 : 0;
    E.IInII[c] = iinii_of(d, E.key, E.dir[c], e);
    d.EtoEdge[3 * k + e] = int(d.edges.size()) - 1;
  }
  auto find = [&](int a, int b) {
    const uint64_t key = key_of(a, b);
    auto it = std::lower_bound(d.edges.begin(), d.edges.end(), key,
                               [](const Edge &E, uint64_t kk) { return E.key < kk; });
    if (it == d.edges.end() || it->key != key) throw SetupError(GOCFD_ERR_MESH, "boundary edge is not an element edge");
    return int(it - d.edges.begin());
  };
  d.EtoE.assign(size_t(3) * K, -1);
  for (const auto &E : d.edges)
    if (E.nconn == 2) {
      d.EtoE[3 * E.tri[0] + E.enr[0]] = E.tri[1];
      d.EtoE[3 * E.tri[1] + E.enr[1]] = E.tri[0];
    }

  // Boundary tags (triangulation.go:52-104)
  bool reversed_last = false;
  for (const auto &t : mesh.tags) {
    const int flag = bc_flag_of_tag(t.tag);
    if (flag < 0) throw SetupError(GOCFD_ERR_MESH, "unable to find BC with base name: [" + t.tag + "]");
    switch (flag) {
      case BC_Far: case BC_IVortex: case BC_Wall: case BC_In: case BC_Out: case BC_Cyl:
        for (auto &e : t.edges) d.edges[find(e.first, e.second)].bc = flag;
        break;
      case BC_PeriodicReversed:
        reversed_last = true;
        [[fallthrough]];
      case BC_Periodic: {
        const size_t l = t.edges.size();
        if (l % 2) throw SetupError(GOCFD_ERR_MESH, "periodic boundaries must be present in pairs");
        std::vector<std::pair<int, int>> c0(t.edges.begin(), t.edges.begin() + l / 2);
        std::vector<std::pair<int, int>> c1(t.edges.begin() + l / 2, t.edges.end());
        c0 = reorder_chain(c0, false);
        c1 = reorder_chain(c1, reversed_last);
        for (size_t i = 0; i < l / 2; ++i) {
          Edge &E1 = d.edges[find(c0[i].first, c0[i].second)];
          Edge &E2 = d.edges[find(c1[i].first, c1[i].second)];
          E1.bc = flag;
          E2.bc = flag;
          if (E1.nconn != 1 || E2.nconn != 1) throw SetupError(GOCFD_ERR_MESH, "periodic edge is not a boundary edge");
          E1.tri[1] = E2.tri[0]; E1.enr[1] = E2.enr[0]; E1.dir[1] = E2.dir[0];
          E1.IInII[1] = iinii_of(d, E2.key, E2.dir[0], E2.enr[0]);
          E1.nconn = 2;
        }
        break;
      }
      default:
        throw SetupError(GOCFD_ERR_MESH, "BC type of tag " + t.tag + " not implemented yet");
    }
  }

  // Geometry (dfr_startup.go:125-175, 256-314; element_utils.go:163-175)
  d.Jdet.resize(K); d.Jinv.resize(size_t(4) * K);
  for (int k = 0; k < K; ++k) {
    const int *v = &d.EToV[3 * k];
    const double xr = 0.5 * (d.VX[v[1]] - d.VX[v[0]]), yr = 0.5 * (d.VY[v[1]] - d.VY[v[0]]);
    const double xs = 0.5 * (d.VX[v[2]] - d.VX[v[0]]), ys = 0.5 * (d.VY[v[2]] - d.VY[v[0]]);
    const double det = xr * ys - xs * yr;
    d.Jdet[k] = det;
    d.Jinv[4 * k] = ys / det; d.Jinv[4 * k + 1] = -xs / det;
    d.Jinv[4 * k + 2] = -yr / det; d.Jinv[4 * k + 3] = xr / det;
  }
  d.FaceNx.assign(size_t(3) * K, 0); d.FaceNy.assign(size_t(3) * K, 0);
  d.IInII.assign(size_t(3) * K, 0); d.FaceEdge.assign(size_t(3) * K, -1);
  auto periodic = [](int bc) { return bc == BC_Periodic || bc == BC_PeriodicReversed; };
  for (size_t ei = 0; ei < d.edges.size(); ++ei) {
    const Edge &E = d.edges[ei];
    for (int c = 0; c < E.nconn; ++c) {
      const int k = E.tri[c], f = E.enr[c];
      const size_t ind = size_t(k) + size_t(K) * f;
      if (c == 1 && periodic(E.bc)) {  // coupled partner face: EdgeNumber -> coupling edge
        d.FaceEdge[ind] = int(ei);
        continue;
      }
      d.IInII[ind] = E.IInII[c];
      if (d.FaceEdge[ind] < 0) d.FaceEdge[ind] = int(ei);
      int a, b;
      oriented(E.key, E.dir[c], a, b);
      const double dx = d.VX[b] - d.VX[a], dy = d.VY[b] - d.VY[a];
      const double oon = 1. / std::sqrt(dx * dx + dy * dy);
      d.FaceNx[ind] = -dy * oon;
      d.FaceNy[ind] = dx * oon;
    }
  }
  d.EdgeLenMax.resize(K);
  for (int k = 0; k < K; ++k) {
    double mx = -1;
    for (int f = 0; f < 3; ++f) mx = std::max(mx, d.IInII[k + size_t(K) * f]);
    d.EdgeLenMax[k] = d.Jdet[k] / mx;
  }
  const int NpInt = d.ops.NpInt, NpFlux = d.ops.NpFlux;
  auto map_pts = [&](const std::vector<double> &R, const std::vector<double> &S, int np,
                     std::vector<double> &X, std::vector<double> &Y) {
    X.resize(size_t(np) * K); Y.resize(size_t(np) * K);
    for (int k = 0; k < K; ++k) {
      const int *v = &d.EToV[3 * k];
      for (int i = 0; i < np; ++i) {
        const double r = R[i], s = S[i];
        X[k + size_t(i) * K] = (-(r + s) * d.VX[v[0]] + (r + 1) * d.VX[v[1]] + (s + 1) * d.VX[v[2]]) * 0.5;
        Y[k + size_t(i) * K] = (-(r + s) * d.VY[v[0]] + (r + 1) * d.VY[v[1]] + (s + 1) * d.VY[v[2]]) * 0.5;
      }
    }
  };
  map_pts(d.ops.solR, d.ops.solS, NpInt, d.SolX, d.SolY);
  map_pts(d.ops.R, d.ops.S, NpFlux, d.FluxX, d.FluxY);
  return d;
}

// ===========================================================================
// Device layout: Hilbert order, tiles, halo lists, face codes
// ===========================================================================
namespace {
// Hilbert index of (x, y) on a 2^16 grid.
uint64_t hilbert(uint32_t x, uint32_t y) {
  uint64_t d = 0;
  for (uint32_t s = 1u << 15; s > 0; s >>= 1) {
    const uint32_t rx = (x & s) ? 1 : 0, ry = (y & s) ? 1 : 0;
    d += uint64_t(s) * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) { x = s - 1 - x; y = s - 1 - y; }
      std::swap(x, y);
    }
  }
  return d;
}
}  // namespace

DeviceLayout build_device_layout(const Discretisation &d, int E, int TP) {
  DeviceLayout L;
  const int K = d.K, NpEdge = d.ops.NpEdge, NpInt = d.ops.NpInt;
  L.K = K;
  L.E = E;
  L.ntiles = (K + E - 1) / E;
  L.Kp = L.ntiles * E;
  const int Kp = L.Kp;
  // Element order along a Hilbert curve through the centroids.
  double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
  std::vector<double> cx(K), cy(K);
  for (int k = 0; k < K; ++k) {
    const int *v = &d.EToV[3 * k];
    cx[k] = (d.VX[v[0]] + d.VX[v[1]] + d.VX[v[2]]) / 3;
    cy[k] = (d.VY[v[0]] + d.VY[v[1]] + d.VY[v[2]]) / 3;
    xmin = std::min(xmin, cx[k]); xmax = std::max(xmax, cx[k]);
    ymin = std::min(ymin, cy[k]); ymax = std::max(ymax, cy[k]);
  }
  const double span = std::max({xmax - xmin, ymax - ymin, 1e-300});
  std::vector<uint64_t> h(K);
  for (int k = 0; k < K; ++k) {
    const uint32_t ix = uint32_t(std::min(65535.0, (cx[k] - xmin) / span * 65535.0));
    const uint32_t iy = uint32_t(std::min(65535.0, (cy[k] - ymin) / span * 65535.0));
    h[k] = hilbert(ix, iy);
  }
  std::vector<int> order(K);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return h[a] < h[b]; });
  L.perm.assign(Kp, -1);
  L.iperm.assign(K, -1);
  for (int kk = 0; kk < K; ++kk) { L.perm[kk] = order[kk]; L.iperm[order[kk]] = kk; }

  // Geometry in internal order (padding: identity element)
  L.Jdet.assign(Kp, 1.0); L.Jinv.assign(size_t(4) * Kp, 0.0);
  L.IInII.assign(size_t(3) * Kp, 1.0); L.Nx.assign(size_t(3) * Kp, 1.0); L.Ny.assign(size_t(3) * Kp, 0.0);
  L.oohK.assign(Kp, 0.0);
  L.hK.assign(Kp, 1.0);
  L.evert.assign(size_t(3) * Kp, 0);
  const double Np12 = double((d.N + 1) * (d.N + 1));
  for (int kk = 0; kk < Kp; ++kk) {
    const int k = L.perm[kk];
    if (k < 0) { L.Jinv[kk] = 1.0; L.Jinv[3 * size_t(Kp) + kk] = 1.0; continue; }
    L.Jdet[kk] = d.Jdet[k];
    for (int c = 0; c < 4; ++c) L.Jinv[size_t(c) * Kp + kk] = d.Jinv[4 * size_t(k) + c];
    const double hK = d.EdgeLenMax[k] / Np12;  // GetHk (dfr_startup.go:89-98)
    L.oohK[kk] = 1. / hK;
    L.hK[kk] = hK;
    for (int c = 0; c < 3; ++c) L.evert[size_t(c) * Kp + kk] = d.EToV[3 * size_t(k) + c];
  }
  // Faces
  L.face_code.assign(size_t(3) * Kp, FK_Inert);
  L.nbr_slot.assign(size_t(3) * Kp, 0);
  L.nbr_elem.assign(size_t(3) * Kp, -1);
  L.nbr_face.assign(size_t(3) * Kp, 0);
  L.Nxo.assign(size_t(3) * Kp, 1.0);
  L.Nyo.assign(size_t(3) * Kp, 0.0);
  L.vooL.assign(size_t(3) * Kp, 0.0);
  L.wsrc.assign(size_t(3) * Kp, 0);
  for (int kk = 0; kk < Kp; ++kk)
    for (int f = 0; f < 3; ++f) L.wsrc[size_t(f) * Kp + kk] = f * Kp + kk;
  std::vector<std::vector<std::pair<int, int>>> halos(L.ntiles);
  std::vector<std::map<std::pair<int, int>, int>> halo_index(L.ntiles);
  for (int kk = 0; kk < K; ++kk) {
    const int k = L.perm[kk], tile = kk / E;
    for (int f = 0; f < 3; ++f) {
      const size_t fi = size_t(f) * Kp + kk, ri = size_t(k) + size_t(K) * f;
      const Edge &Ed = d.edges[d.FaceEdge[ri]];
      const int side = (Ed.tri[0] == k && Ed.enr[0] == f) ? 0 : 1;
      const Edge &own = d.edges[d.EtoEdge[3 * size_t(k) + f]];
      int code = 0;
      if (own.tri[0] == k && own.enr[0] == f) code |= FC_WAVE;
      else  // the edge's speed is taken on its owner's face
        L.wsrc[fi] = own.enr[0] * Kp + L.iperm[own.tri[0]];
      L.IInII[fi] = d.IInII[ri];
      L.Nxo[fi] = d.FaceNx[ri];
      L.Nyo[fi] = d.FaceNy[ri];
      // normal used by the Riemann solve: the flux owner's
      const size_t oi = size_t(Ed.tri[0]) + size_t(K) * Ed.enr[0];
      if (Ed.nconn == 2) {
        L.Nx[fi] = d.FaceNx[oi];
        L.Ny[fi] = d.FaceNy[oi];
        if (side == 0) code |= FC_FLUX_OWNER;
        const int pk = Ed.tri[1 - side], pf = Ed.enr[1 - side];
        const int pkk = L.iperm[pk];
        L.nbr_elem[fi] = pkk;
        L.nbr_face[fi] = pf;
        // StoreEdgeViscousFlux's edge length (edges.go:195-201): owner face 1
        // is the hypotenuse of the unit triangle
        const double len = Ed.enr[0] == 1 ? (2. * std::sqrt(2.)) * Ed.IInII[0] : 2. * Ed.IInII[0];
        L.vooL[fi] = 1. / len;
        if (pkk / E == tile) {
          L.nbr_slot[fi] = pf * NpEdge * TP + (pkk - tile * E);
        } else {
          auto key = std::make_pair(pkk, pf);
          auto it = halo_index[tile].find(key);
          int hidx;
          if (it == halo_index[tile].end()) {
            hidx = int(halos[tile].size());
            halos[tile].push_back(key);
            halo_index[tile][key] = hidx;
          } else {
            hidx = it->second;
          }
          L.nbr_slot[fi] = int(0x80000000u | uint32_t(3 * NpEdge * TP + hidx * NpEdge));
        }
        code |= FK_Shared;
      } else {
        L.Nx[fi] = d.FaceNx[ri];
        L.Ny[fi] = d.FaceNy[ri];
        code |= FC_FLUX_OWNER;
        switch (Ed.bc) {
          case BC_None: code |= FK_Open; break;
          // bits 8.. select the far state: FSFar, FSIn, FSOut
          case BC_Far: code |= FK_Far; break;
          case BC_In: code |= FK_Far | (1 << 8); break;
          case BC_Out: code |= FK_Far | (2 << 8); break;
          case BC_IVortex: {
            code |= FK_IVortex;
            code |= (L.nbfaces << 8);
            for (int p = 0; p < NpEdge; ++p) {
              const size_t pi = size_t(k) + size_t(2 * NpInt + f * NpEdge + p) * K;
              L.bpoint_xy.push_back(d.FluxX[pi]);
              L.bpoint_xy.push_back(d.FluxY[pi]);
            }
            ++L.nbfaces;
            break;
          }
          case BC_Wall: case BC_Cyl: code |= FK_Wall; break;
          default: code |= FK_Inert; break;  // uncoupled periodic: zero flux (edges.go:436-438)
        }
      }
      L.face_code[fi] = code;
    }
  }
  // Shared faces with Riemann work, per tile
  // (cross-tile faces first: item i of tile t receives its neighbour traces
  // in halo-trace slot t * Fmax + i, written by the neighbour's tile)
  std::vector<int> item_of(size_t(3) * Kp, -1);  // flux-list item of each cross-tile face
  {
    std::vector<std::vector<int>> fl(L.ntiles);
    std::vector<std::vector<std::array<double, 4>>> fg(L.ntiles);
    for (int pass = 0; pass < 2; ++pass)
    for (int kk = 0; kk < K; ++kk)
      for (int f = 0; f < 3; ++f) {
        const size_t fi = size_t(f) * Kp + kk;
        const int code = L.face_code[fi];
        if ((code & 15) != FK_Shared) continue;
        const bool in_tile = !(uint32_t(L.nbr_slot[fi]) >> 31);
        const bool owner = code & FC_FLUX_OWNER;
        if (in_tile && !owner) continue;  // the in-tile owner fills this slot
        if (in_tile != (pass == 1)) continue;
        if (!in_tile) item_of[fi] = int(fl[kk / E].size());
        const int base = int(uint32_t(L.nbr_slot[fi]) & 0x7fffffffu);
        if (base >= (1 << 18)) throw SetupError(GOCFD_ERR_MESH, "tile halo too large for the flux-list encoding");
        fl[kk / E].push_back((kk % E) | (f << 8) | ((in_tile && owner) ? 1 << 10 : 0) | (in_tile ? 0 : 1 << 11) |
                             (owner ? 1 << 12 : 0) | (base << 13));
        double w2 = 0.0;
        if (in_tile) {  // the partner slot pf * NpEdge * TP + e2
          const int pf = base / (NpEdge * TP), e2 = base - pf * NpEdge * TP;
          w2 = -L.IInII[size_t(pf) * Kp + (kk / E) * E + e2];
        }
        fg[kk / E].push_back({L.Nx[fi], L.Ny[fi], owner ? L.IInII[fi] : -L.IInII[fi], w2});
      }
    L.Fmax = 1;
    for (auto &v : fl) L.Fmax = std::max<int>(L.Fmax, int(v.size()));
    L.flux_list.assign(size_t(L.ntiles) * L.Fmax, 0);
    L.flux_geo.assign(size_t(L.ntiles) * L.Fmax * 4, 0.0);
    L.flux_cnt.assign(L.ntiles, 0);
    for (int t = 0; t < L.ntiles; ++t) {
      L.flux_cnt[t] = int(fl[t].size());
      std::copy(fl[t].begin(), fl[t].end(), L.flux_list.begin() + size_t(t) * L.Fmax);
      for (size_t i = 0; i < fg[t].size(); ++i)
        for (int c = 0; c < 4; ++c) L.flux_geo[(size_t(t) * L.Fmax + i) * 4 + c] = fg[t][i][c];
    }
  }
  // Cross-tile faces per tile, with the neighbour face each one exports to
  {
    std::vector<std::vector<std::pair<int, int>>> xp(L.ntiles);
    for (int kk = 0; kk < K; ++kk)
      for (int f = 0; f < 3; ++f) {
        const size_t fi = size_t(f) * Kp + kk;
        if ((L.face_code[fi] & 15) != FK_Shared || !(uint32_t(L.nbr_slot[fi]) >> 31)) continue;
        // destination: the item of the neighbour's face in its tile's list
        const size_t ni = size_t(L.nbr_face[fi]) * Kp + L.nbr_elem[fi];
        if (item_of[ni] < 0) throw SetupError(GOCFD_ERR_MESH, "cross-tile face without a flux-list item");
        xp[kk / E].push_back({(kk % E) | (f << 8), (L.nbr_elem[fi] / E) * L.Fmax + item_of[ni]});
      }
    L.Xmax = 1;
    for (auto &v : xp) L.Xmax = std::max<int>(L.Xmax, int(v.size()));
    L.xp_list.assign(size_t(L.ntiles) * L.Xmax, 0);
    L.xp_dst.assign(size_t(L.ntiles) * L.Xmax, 0);
    L.xp_cnt.assign(L.ntiles, 0);
    for (int t = 0; t < L.ntiles; ++t) {
      L.xp_cnt[t] = int(xp[t].size());
      for (size_t i = 0; i < xp[t].size(); ++i) {
        L.xp_list[size_t(t) * L.Xmax + i] = xp[t][i].first;
        L.xp_dst[size_t(t) * L.Xmax + i] = xp[t][i].second;
      }
    }
  }
  // Non-shared faces per tile (the boundary loop of the stage kernel)
  L.bnd_cnt.assign(L.ntiles, 0);
  for (int kk = 0; kk < Kp; ++kk)
    for (int f = 0; f < 3; ++f)
      if ((L.face_code[size_t(f) * Kp + kk] & 15) != FK_Shared) ++L.bnd_cnt[kk / E];
  // Faces whose owner-side trace enters the CFL wave-speed maximum, per tile
  std::vector<std::vector<int>> waves(L.ntiles);
  for (int kk = 0; kk < K; ++kk)
    for (int f = 0; f < 3; ++f)
      if (L.face_code[size_t(f) * Kp + kk] & FC_WAVE) waves[kk / E].push_back((kk % E) | (f << 8));
  L.Wmax = 1;
  for (auto &w : waves) L.Wmax = std::max<int>(L.Wmax, int(w.size()));
  L.wave_list.assign(size_t(L.ntiles) * L.Wmax, 0);
  L.wave_code.assign(size_t(L.ntiles) * L.Wmax, 0);
  L.wave_cnt.assign(L.ntiles, 0);
  for (int t = 0; t < L.ntiles; ++t) {
    L.wave_cnt[t] = int(waves[t].size());
    std::copy(waves[t].begin(), waves[t].end(), L.wave_list.begin() + size_t(t) * L.Wmax);
    for (size_t i = 0; i < waves[t].size(); ++i) {
      const int w = waves[t][i];
      L.wave_code[size_t(t) * L.Wmax + i] = L.face_code[size_t(w >> 8) * Kp + size_t(t) * E + (w & 255)];
    }
  }
  L.Hmax = 1;
  for (auto &hl : halos) L.Hmax = std::max<int>(L.Hmax, int(hl.size()));
  L.halo_elem.assign(size_t(L.ntiles) * L.Hmax, -1);
  L.halo_face.assign(size_t(L.ntiles) * L.Hmax, 0);
  for (int t = 0; t < L.ntiles; ++t)
    for (size_t i = 0; i < halos[t].size(); ++i) {
      L.halo_elem[size_t(t) * L.Hmax + i] = halos[t][i].first;
      L.halo_face[size_t(t) * L.Hmax + i] = halos[t][i].second;
    }
  return L;
}

}  // namespace gocfd

namespace gocfd {

// NewAliasShockFinder (DG2D/dfr_shock_capturing.go:71-107), ModalFilter with
// RecommendedFilterParameters (:43-69, :109-140), createInterpolationStencil
// (model_problems/Euler2D/dissipation.go:414-449)
std::vector<double> build_shock_operators(const Operators &o) {
  const int N = o.N, Np = o.NpInt, NpEdge = o.NpEdge, NpFlux = o.NpFlux, NR = Np + 3 * NpEdge;
  const size_t n2 = size_t(Np) * Np;
  std::vector<double> out(5 * n2 + Np + size_t(NR) * NpFlux + 3 * size_t(NR), 0.0);
  double *P = out.data(), *M = P + n2, *D = M + n2, *V = D + n2, *Vi = V + n2, *mf = Vi + n2;
  double *Div = mf + Np, *Bary = Div + size_t(NR) * NpFlux;
  // Clip = V diag(modes with i+j < N) Vinv, D = I - Clip, P = Mass D
  std::vector<double> diag(Np), tmp(n2), clip(n2, 0.0);
  int ii = 0;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N - i; ++j) diag[ii++] = (i + j >= N) ? 0.0 : 1.0;
  for (int i = 0; i < Np; ++i)
    for (int j = 0; j < Np; ++j) tmp[size_t(i) * Np + j] = o.V[size_t(i) * Np + j] * diag[j];
  for (int i = 0; i < Np; ++i)
    for (int j = 0; j < Np; ++j) {
      double acc = 0;
      for (int l = 0; l < Np; ++l) acc += tmp[size_t(i) * Np + l] * o.Vinv[size_t(l) * Np + j];
      clip[size_t(i) * Np + j] = acc;
    }
  for (int i = 0; i < Np; ++i)
    for (int j = 0; j < Np; ++j) D[size_t(i) * Np + j] = (i == j ? 1.0 : 0.0) - clip[size_t(i) * Np + j];
  for (int i = 0; i < Np; ++i)
    for (int j = 0; j < Np; ++j) {
      double acc = 0;
      for (int l = 0; l < Np; ++l) acc += o.Mass[size_t(i) * Np + l] * D[size_t(l) * Np + j];
      P[size_t(i) * Np + j] = acc;
    }
  std::copy(o.Mass.begin(), o.Mass.end(), M);
  std::copy(o.V.begin(), o.V.end(), V);
  std::copy(o.Vinv.begin(), o.Vinv.end(), Vi);
  double alpha;
  int sp;
  if (N <= 2) { alpha = 4.0; sp = 4; }
  else if (N == 3) { alpha = 6.0; sp = 4; }
  else if (N == 4) { alpha = 8.0; sp = 5; }
  else if (N == 5) { alpha = 10.0; sp = 6; }
  else if (N == 6) { alpha = 14.0; sp = 6; }
  else { alpha = 18.0; sp = 7; }
  ii = 0;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N - i; ++j) {
      const int p = i + j;
      mf[ii++] = (N <= 1 || p == 0) ? 1.0 : std::exp(-alpha * std::pow(double(p) / double(N), double(sp)));
    }
  // rows of Div at the used RT points, and their barycentric coordinates
  // with respect to the vertices (-1,-1), (1,-1), (-1,1)
  for (int r = 0; r < NR; ++r) {
    const int row = r < Np ? r : r + Np;
    for (int j = 0; j < NpFlux; ++j) Div[size_t(r) * NpFlux + j] = o.Div[size_t(row) * NpFlux + j];
    const double rr = o.R[row], ss = o.S[row];
    Bary[3 * r] = -0.5 * (rr + ss);
    Bary[3 * r + 1] = 0.5 * (rr + 1.0);
    Bary[3 * r + 2] = 0.5 * (ss + 1.0);
  }
  return out;
}

}  // namespace gocfd
