// Host-side setup of the 2D DFR discretisation: mesh ingest, edge topology,
// Raviart-Thomas / Lagrange operators, element geometry, and the device
// ordering (Hilbert-curve element order, tiles, halo lists).
//
// Mirrors the reference's DG2D.NewDFR2D (DG2D/dfr_startup.go:43-82) and
// readfiles (readSU2Grid.go, readGambitGrid.go); see each function.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace gocfd {

struct SetupError : std::runtime_error {
  int code;
  SetupError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Boundary flags, numbered as types.BCFLAG (types/cfd.go:13-26).
enum BCFlag : int {
  BC_None = 0, BC_In, BC_Dirichlet, BC_Slip, BC_Far, BC_Wall, BC_Cyl, BC_Neuman, BC_Out,
  BC_IVortex, BC_Periodic, BC_PeriodicReversed
};

struct BoundaryTag {
  std::string tag;                          // lower-cased, trimmed (types.NewBCTAG)
  std::vector<std::pair<int, int>> edges;   // ordered vertex pairs (types.EdgeInt)
};

struct Mesh {
  int K = 0, Nv = 0;
  std::vector<double> VX, VY;
  std::vector<int> EToV;                    // K x 3
  std::vector<BoundaryTag> tags;            // file order; repeated tags appended
  void add_tag(const std::string &tag, std::vector<std::pair<int, int>> e);
};

Mesh read_mesh(const std::string &path);
int bc_flag_of_tag(const std::string &tag);  // -1 if unknown

// Reference-element operators for solution order N (RT order N+1).
struct Operators {
  int N = 0, NpInt = 0, NpEdge = 0, NpFlux = 0;
  std::vector<double> R, S;          // RT point locations [NpFlux]
  std::vector<double> Div;           // [NpFlux x NpFlux]
  std::vector<double> DivInt;        // [NpInt x NpFlux]
  std::vector<double> Interp;        // FluxEdgeInterp [3NpEdge x NpInt]
  std::vector<double> solR, solS, solW;
  std::vector<double> V, Vinv, Mass; // solution element [NpInt x NpInt]
};
Operators build_operators(int N);
// The solution basis interpolated to arbitrary reference points (r, s):
// [npts x NpInt] = V(r, s) Vinv (JacobiBasis2D.GetInterpMatrix,
// DG2D/basis_polynomials.go:49-53)
std::vector<double> interp_matrix(const Operators &o, int npts, const double *r, const double *s);

struct Edge {
  uint64_t key = 0;
  int nconn = 0;
  int tri[2] = {-1, -1}, enr[2] = {-1, -1}, dir[2] = {-1, -1};
  int bc = BC_None;
  double IInII[2] = {0, 0};
};

// Topology and geometry in the reference element numbering.
struct Discretisation {
  int N = 0, K = 0, Nv = 0;
  Operators ops;
  std::vector<double> VX, VY;
  std::vector<int> EToV;
  std::vector<Edge> edges;          // sorted by key
  std::vector<int> EtoE;            // K x 3
  std::vector<int> EtoEdge;         // K x 3 (own edge index: EtoEdges)
  std::vector<int> FaceEdge;        // 3 x K (EdgeNumber as an edge index)
  std::vector<double> Jdet, Jinv;   // K, K x 4
  std::vector<double> FaceNx, FaceNy, IInII;  // 3 x K
  std::vector<double> EdgeLenMax;   // K
  std::vector<double> SolX, SolY;   // NpInt x K
  std::vector<double> FluxX, FluxY; // NpFlux x K
};
Discretisation build_discretisation(const Mesh &mesh, int N);

// Face kinds seen by the device kernels.
enum FaceKind : int {
  FK_Shared = 0,   // flux against a neighbour trace (interior or periodic)
  FK_Open = 1,     // unconnected, no BC tag: physical flux of the own trace
  FK_Far = 2,      // Riemann invariant far field; code >> 8: 0 Far, 1 In, 2 Out
  FK_IVortex = 3,  // analytic isentropic vortex + Riemann BC
  FK_Wall = 4,     // pressure-only wall flux (Wall, Cyl)
  FK_Inert = 5     // padding element face
};
constexpr int FC_FLUX_OWNER = 1 << 4;  // computes F*(own, nbr, n_owner)
constexpr int FC_WAVE = 1 << 5;        // contributes to the CFL wave-speed max

// Device ordering: element permutation, tiles and halo lists.
struct DeviceLayout {
  int K = 0, E = 0, ntiles = 0, Kp = 0, Hmax = 0;
  std::vector<int> perm;            // perm[kk] = reference k (or -1 for padding)
  std::vector<int> iperm;           // iperm[k] = kk
  // per internal element kk, per face f (index f*Kp + kk)
  std::vector<int> face_code;       // kind | flags | (boundary-point index << 8)
  std::vector<int> nbr_slot;        // tile-relative LDS trace slot of nbr point 0; bit 31: halo
  std::vector<int> nbr_elem;        // internal index of the neighbour (or -1)
  std::vector<double> Jdet, Jinv, IInII, Nx, Ny, oohK;  // geometry in internal order
  std::vector<int> halo_elem, halo_face;                // [ntiles x Hmax], -1 unused
  int Wmax = 0;                                         // owned (wave-speed) faces per tile, max
  std::vector<int> wave_list;                           // [ntiles x Wmax]: e | face << 8
  std::vector<int> wave_cnt;                            // [ntiles]
  std::vector<int> wave_code;                           // [ntiles x Wmax]: face_code of each wave_list face
  // Shared faces whose Riemann flux this tile evaluates: every face with an
  // out-of-tile neighbour, and the owner side of in-tile pairs (which also
  // fills the partner's slot). [ntiles x Fmax]: e | face << 8 | partner << 10
  // | (neighbour out of tile) << 11 | (flux owner) << 12 | nbr_slot base << 13
  int Fmax = 0;
  std::vector<int> flux_list;
  // per item [ntiles x Fmax][4]: the owner normal (nx, ny), this side's
  // IInII signed by the side (+ owner, - other side) and the in-tile
  // partner's IInII negated (0 without partner)
  std::vector<double> flux_geo;
  std::vector<int> flux_cnt;                            // [ntiles]
  // Faces of the tile's elements (padding included) that are not shared:
  // boundary, inert and padding faces, whose flux the boundary loop sets.
  // Tiles with none skip that loop. [ntiles]
  std::vector<int> bnd_cnt;
  // Faces whose neighbour lies in another tile, per tile (the export list of
  // the halo traces): [ntiles x Xmax] e | face << 8, the halo-trace item slot
  // of the neighbour's face (its tile * Fmax + its flux-list item; the
  // cross-tile faces lead each tile's flux list), and [ntiles] counts
  int Xmax = 0;
  std::vector<int> xp_list, xp_dst, xp_cnt;
  std::vector<double> bpoint_xy;                        // IVortex boundary points [nb][NpEdge][2]
  int nbfaces = 0;
  // shock capturing and local time stepping, per face [3][Kp] / element [Kp]
  std::vector<int> nbr_face;        // neighbour's face number (or 0)
  std::vector<double> Nxo, Nyo;     // own outward normals
  std::vector<double> vooL;         // 1 / owner edge length, (2 or 2 sqrt 2) IInII (shared faces)
  std::vector<int> wsrc;            // face (f*Kp + kk) whose wave speed bounds this face's edge
  std::vector<int> evert;           // [3][Kp] vertex ids
  std::vector<double> hK;           // EdgeLenMax / (N+1)^2
};
DeviceLayout build_device_layout(const Discretisation &d, int tile_elems, int trace_pitch);

// Operator block of the shock-capturing kernels for order N (layout of
// gocfd::dev::ShockArgs::ops): P = Mass (I - Clip), Mass, D = I - Clip,
// V, Vinv [NpInt^2] each, modal filter factors [NpInt], the rows of Div used
// (0..NpInt-1 and 2NpInt..NpFlux-1) [NR][NpFlux], and the barycentric
// coordinates of those RT points [NR][3]; NR = NpInt + 3 NpEdge.
std::vector<double> build_shock_operators(const Operators &ops);

}  // namespace gocfd
