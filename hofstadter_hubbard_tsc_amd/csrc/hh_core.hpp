// Host-side core of the MI355X SU(2) DMRG inner loop: SU(2) coefficients,
// lattice geometry, block / operator layouts and the task compilers that
// lower every step of the reference inner loop to grouped complex GEMM tiles.
#pragma once
#include <cstdint>
#include <cstddef>
#include <complex>
#include <map>
#include <string>
#include <vector>

namespace hh {

using cplx = std::complex<double>;

// ----------------------------------------------------------------------------
// SU(2) coefficients (doubled-integer arguments)
// ----------------------------------------------------------------------------
// Wigner 6j {j1 j2 j3; l1 l2 l3}, Racah formula as in W6JS
// (code/update.f90:1665-1716), evaluated in long double.
double w6j(int j1, int j2, int j3, int l1, int l2, int l3);
inline double phase2(int two_x) { return ((two_x / 2) & 1) ? -1.0 : 1.0; }  // (-1)^x, 2x even
// <(b' s)J'||T(1)||(b s)J> / <b'||T||b>    (Edmonds 7.1.7)
double e1(int b2p, int b2, int s2, int J2p, int J2, int k2);
// <(b s')J'||U(2)||(b s)J> / <s'||U||s>    (Edmonds 7.1.8)
double e2(int s2p, int s2, int b2, int J2p, int J2, int k2);
// matrix factor of [T(1) x U(2)]^0 on (b s)J  (reduced / sqrt(2J+1))
double x12(int b2p, int b2, int s2p, int s2, int J2, int k2);

// ----------------------------------------------------------------------------
// model, lattice and geometry
// ----------------------------------------------------------------------------
enum ModelKind { HUBBARD = 0, TJ = 1 };
enum LatticeKind { SQUARE = 0, TRIANGULAR = 1 };
// bond phases: the Landau gauge of `flux`, or get_phase's pattern with
// gauge = 1 and x-bond angle `theta` (triangular lattice)
enum GaugeKind { GAUGE_LANDAU = 0, GAUGE_REFERENCE = 1 };

struct ModelParams {
  int kind = HUBBARD;
  int Lx = 4, Ly = 2;
  int lattice = SQUARE;
  int periodic_y = 1;
  double t = 1.0, U = 8.0, J = 0.4, flux = 0.0;
  int n_elec = 8;
  int gauge = GAUGE_LANDAU;
  double theta = 0.0;
  int nsite() const { return Lx * Ly; }
};

struct Coupling {        // one bond between two positions
  cplx hop;              // amplitude of c^dag_u c_v (u < v)
  double J, V;           // S.S and n n couplings
};

// Superblock geometry of one step (``dmrg_setup``, code/warm_up.f90:347-376).
struct Geometry {
  int ns = 0, ne = 0, ncol = 0, totsite = 0;
  std::map<std::pair<int, int>, Coupling> bonds;   // key (u<v)
  int sys_pos(int k) const { return k; }
  int env_pos(int k) const { return totsite + 1 - k; }
  int sys_centre() const { return ns + 1; }
  int env_centre() const { return totsite - ne; }
  // coupling for (pu, pv) in that orientation (hop is amp of c^dag_pu c_pv)
  bool coupling(int pu, int pv, Coupling& out) const;
};
Geometry make_geometry(const ModelParams& m, int ns, int ne);
std::vector<int> boundary_sites(const ModelParams& m, int length);
int electrons_for_length(const ModelParams& m, int nsites);

// ----------------------------------------------------------------------------
// sectors, enlarged bases, super bases
// ----------------------------------------------------------------------------
struct Sector { int N, S2, dim; };

struct SubSector { int b, s, pos, dim; };       // block sector, site sector
struct EnlBasis {
  std::vector<Sector> sec;                      // enlarged sectors (Get_basis order)
  std::vector<std::vector<SubSector>> subs;
  std::vector<Sector> bl, st;                   // block and site sectors
  int find(int N, int S2) const;
  int total_dim() const;
};
EnlBasis enlarge_basis(const std::vector<Sector>& bl, const std::vector<Sector>& st);

struct SuperBasis {
  std::vector<std::pair<int, int>> pairs;       // (sys enlarged sector, env enlarged sector)
  std::vector<int64_t> off;                     // element offset of each psi block
  std::vector<int> rows, cols;
  int64_t dim = 0;
};
SuperBasis super_basis(const EnlBasis& s, const EnlBasis& e, int N);

std::vector<Sector> site_sectors(int kind);
// reduced element <s'||O||s> of the site operator `kind` (OP_CDAG, OP_S,
// OP_N) between sectors sp and s of site_sectors (Get_site_operator_su2,
// code/warm_up.f90:818-991)
cplx site_val(int model_kind, int kind, int sp, int s);
// annihilator partner of the site c^dag over the site sectors st:
// <j||a~||j'> = -(-1)^(1/2+j'-j) conj(<j'||c^dag||j>)
cplx site_tilde(int model_kind, const std::vector<Sector>& st, int sp, int s);

// ----------------------------------------------------------------------------
// operator layouts (one per operator type on a block)
// ----------------------------------------------------------------------------
struct SubBlk { int bra, ket, rows, cols; int64_t off; };
struct OpLayout {
  int k2 = 0, dN = 0;
  std::vector<SubBlk> subs;
  int64_t size = 0;                             // complex elements
  std::map<std::pair<int, int>, int> index;     // (bra, ket) -> position in subs
  int find(int bra, int ket) const {
    auto it = index.find({bra, ket});
    return it == index.end() ? -1 : it->second;
  }
};
// All allowed (bra <- ket) sub-blocks of a rank-k2/2, dN tensor on sectors.
OpLayout make_layout(const std::vector<Sector>& sec, int k2, int dN);

// Operator kinds stored per boundary site: 0 = c^dag, 1 = S, 2 = n.
// OP_PAIR (never stored in a block) is the singlet pair creator [c^dag x
// c^dag]^0 of two block sites: rank 0, N + 2, bosonic.
enum OpKind { OP_CDAG = 0, OP_S = 1, OP_N = 2, OP_PAIR = 3 };
inline int op_k2(int k) { return k == OP_CDAG ? 1 : (k == OP_S ? 2 : 0); }
inline int op_dN(int k) { return k == OP_CDAG ? 1 : (k == OP_PAIR ? 2 : 0); }
inline bool op_fermion(int k) { return k == OP_CDAG; }
int n_op_kinds(int model_kind);    // Hubbard: c^dag; t-J: c^dag, S, n

// Host description of a block whose data lives in one device arena:
// H (one dense reduced matrix per sector, column-major) followed by the
// boundary operators, site by site, kind by kind.
struct BlockDesc {
  int side = 0;                                 // 0 sys, 1 env
  int length = 0;
  std::vector<Sector> sec;
  std::vector<int64_t> h_off;                   // per sector
  OpLayout lay[3];                              // per op kind
  std::vector<int> sites;                       // boundary sites (block index)
  std::vector<std::vector<int64_t>> op_off;     // [site idx][kind] -> arena offset
  int64_t size = 0;                             // arena size (complex elements)
};
// Lay out H and the boundary operators of a block with the given sectors.
void layout_block(BlockDesc& b, int nkinds);

// ----------------------------------------------------------------------------
// grouped-GEMM task format (shared with the HIP engine)
// ----------------------------------------------------------------------------
enum Op : uint8_t { OP_NN = 0, OP_T = 1, OP_C = 2, OP_R = 3 };  // N, transpose, conj-transpose, conj
// One contribution to an output tile: C += coef * op(A) * op(B) (K > 0), or
// C += coef * op(A) elementwise (K == 0, A is M x N).
struct Term {
  double cr, ci;
  int64_t a_off, b_off;
  int32_t a_ld, b_ld;
  int32_t K;
  uint8_t a_buf, b_buf, a_op, b_op;
};
struct Tile {
  int64_t c_off;
  int32_t c_ld, M, N;
  int32_t t0, t1;          // term range
  int32_t c_buf, accumulate;
};
static_assert(sizeof(Term) == 48, "Term layout");
static_assert(sizeof(Tile) == 36 || sizeof(Tile) == 40, "Tile layout");

struct TaskList {
  std::vector<Tile> tiles;
  std::vector<Term> terms;
  double flops_alg = 0.0;       // reference-order algorithmic flops
  double flops_exec = 0.0;      // flops the engine executes
  void add_tile(int buf, int64_t off, int ld, int M, int N, int acc = 0) {
    Tile t{off, ld, M, N, (int32_t)terms.size(), (int32_t)terms.size(), buf, acc};
    tiles.push_back(t);
  }
  void add_term(const Term& t) {
    terms.push_back(t);
    tiles.back().t1 = (int32_t)terms.size();
  }
};

// buffer slots of a task list
enum Buf : uint8_t {
  B_IN = 0, B_OUT = 1, B_SYS = 2, B_ENV = 3, B_X = 4, B_MRG = 5, B_AUX = 6, B_AUX2 = 7
};

// ----------------------------------------------------------------------------
// matvec compiler (hmerge + model_wave)
// ----------------------------------------------------------------------------
// Merged operator descriptions (hmerge): each is a linear combination of
// stored block operators of one kind; materialised on the device.
struct MergeSpec {
  int side;          // 0 sys block, 1 env block
  int kind;          // OP_CDAG / OP_S / OP_N
  std::vector<std::pair<int, cplx>> parts;      // (site index in block.sites, coefficient)
  int64_t off = 0;   // offset in the merged arena
};

struct MatvecPlan {
  SuperBasis sb;
  EnlBasis sbs, ebs;
  std::vector<MergeSpec> merges;
  int64_t merged_size = 0;
  int64_t x_size = 0;            // stage-1 scratch
  TaskList stage1, stage2;       // X = A psi ; out = sum of terms
  TaskList merge_tasks;          // builds merged operators (K==0 axpy tiles)
  double bytes_alg = 0.0;        // 16*(2 D) + 16*(unique operator sub-block elements)
};

// One operator of a two-point measurement term: a block operator (layout
// `lay` over the block's sectors, data at `base` of slot `buf`) or the centre
// site's operator of that kind.
struct MeasurePiece {
  bool block = false;
  bool tilde = false;   // env c^dag layout read as the annihilator partner (as in the hopping bond)
  int kind = OP_S;
  uint8_t buf = B_SYS;
  int64_t base = 0;
  const OpLayout* lay = nullptr;
};
// [A_sys (x) B_env]^0 scaled by amp, the coupling form of the H.psi cross terms
struct MeasureTerm {
  MeasurePiece s, e;
  cplx amp = 1.0;
  // within = true: [A_block (x) B_site]^0 inside the enlarged block of
  // `side` (s is the block operator, e that side's centre site operator),
  // the form of the block-site bond of H.psi
  bool within = false;
  int side = 0;
};
// With only != nullptr the plan holds that one term instead of the model's H
// (block Hamiltonians, U and every bond left out); sys.sec / env.sec give the
// bases and the blocks' own operators are not read.
MatvecPlan compile_matvec(const ModelParams& m, const Geometry& g, const BlockDesc& sys,
                          const BlockDesc& env, int N, const MeasureTerm* only = nullptr);

// ----------------------------------------------------------------------------
// growth + renormalisation compiler (Get_Hamiltonian_Sys/Env + update_trun_*)
// ----------------------------------------------------------------------------
struct KeptSector { int a; int keep; int64_t u_off; int u_ld; };  // U = columns of eigvecs
struct GrowPlan {
  EnlBasis eb;
  BlockDesc out;                 // layout of the new block
  std::vector<MergeSpec> merges; // merged block ops for the new-site bonds
  int64_t merged_size = 0;
  int64_t x_size = 0;
  TaskList merge_tasks, stage1, stage2;
};
// U buffers live in slot B_AUX; the old block in B_SYS, the new block in B_OUT.
GrowPlan compile_grow(const ModelParams& m, const Geometry& g, const BlockDesc& blk,
                      const std::vector<KeptSector>& kept);

// ----------------------------------------------------------------------------
// density matrix and wave transformation compilers
// ----------------------------------------------------------------------------
// rho blocks in super-basis order; slot B_IN = psi, output B_OUT.
TaskList compile_density(const SuperBasis& sb, int side, std::vector<int>& rho_sector,
                         std::vector<int64_t>& rho_off, std::vector<int>& rho_dim);

// LR: old psi (B_IN) -> new psi (B_OUT); sys U in B_AUX, env U in B_AUX2.
struct TransPlan { TaskList stage1, stage2; int64_t x_size = 0; };
TransPlan compile_wave_trans_lr(const SuperBasis& old_sb, const EnlBasis& old_sys, const EnlBasis& old_env,
                                const std::vector<KeptSector>& sys_trun,
                                const std::vector<KeptSector>& env_trun_of_E,
                                const EnlBasis& new_sys, const EnlBasis& new_env, const SuperBasis& new_sb);
// ----------------------------------------------------------------------------
// site densities (measure_operator_dia, code/measure.f90:3-61)
// ----------------------------------------------------------------------------
// Rank-0 operators of a block are stored sector-diagonal: the d x d block of
// every sector, in sector order, one operator after another.
int64_t diag_op_size(const std::vector<Sector>& sec);
std::vector<int64_t> diag_op_offsets(const std::vector<Sector>& sec);
struct TrackPlan {
  EnlBasis eb;
  int64_t x_size = 0, old_size = 0, new_size = 0;
  TaskList stage1, stage2;
};
// One growth step of the operator chain of Get_sys_oper_dia / Get_env_oper_dia:
// the nold tracked operators of block `prev` (slot B_SYS) embedded in the
// enlarged basis and renormalised with the truncation `kept` (U in B_AUX),
// plus the density of the new site; nold + 1 operators of the new block to
// B_OUT, scratch in B_X.
TrackPlan compile_track_step(const ModelParams& m, const std::vector<Sector>& prev,
                             const std::vector<KeptSector>& kept, int nold);
// The same step for an explicit list of new operators: CARRY k embeds old
// operator k, SITE is the density of the new site (k = 0) or its double
// occupancy n_up n_down (k = 1), PRODUCT k is old operator k times the new
// site's density (the two-site operator n_k n_new of Get_sys2b, created when
// its second site joins the block).  New operator q goes to B_OUT at
// q * new_size.
// One step of the chain of a rank-1 operator (the spin S of every block
// site): the nold operators of block `prev` (layout make_layout(prev, 2, 0),
// slot B_SYS, old_size apart) embedded in the enlarged basis
// (update_block_ndia_*) plus the new site's S (update_site_ndia_*), each
// renormalised as U_bra^dag O U_ket; nold + 1 operators in
// make_layout(new sectors, 2, 0) to B_OUT, new_size apart.
TrackPlan compile_track_spin(const ModelParams& m, const std::vector<Sector>& prev,
                             const std::vector<KeptSector>& kept, int nold);
// The same chain for an operator kind of any rank (OP_CDAG: c^dag, rank 1/2,
// N + 1, fermionic, with the fermion signs of update_block_ndia_* /
// update_site_ndia_* for `side`), layout make_layout(sectors, k2, dN).
TrackPlan compile_track_rank(const ModelParams& m, const std::vector<Sector>& prev,
                             const std::vector<KeptSector>& kept, int nold, int kind, int side);
// The same step carrying only the old operators `carry` (in that order) and,
// when add_site, the new site's operator after them.
TrackPlan compile_track_rank(const ModelParams& m, const std::vector<Sector>& prev,
                             const std::vector<KeptSector>& kept, const std::vector<int>& carry, bool add_site,
                             int kind, int side);
// One step of the chain of singlet pair creators P_12 = [c^dag_1 x
// c^dag_2]^0 of block sites (Get_sys_two, code/correlation.f90:331-367):
// CARRY q embeds old pair operator q (slot B_SYS, layout make_layout(prev, 0,
// 2)); CREATE k forms the pair of old carried c^dag number k (slot B_ENV,
// layout make_layout(prev, 1, 1), cd_stride apart) with the new site's
// c^dag.  Each is renormalised as U_bra^dag P U_ket into make_layout(new, 0,
// 2) at q * new_size of B_OUT.
struct PairSrc {
  enum What { CARRY, CREATE } what;
  int k;
};
TrackPlan compile_track_pairs(const ModelParams& m, const std::vector<Sector>& prev,
                              const std::vector<KeptSector>& kept, const std::vector<PairSrc>& srcs, int side);
// SPIN_PRODUCT k is S_k . S_new = -sqrt(3) [S_k x S_new]^0 of carried spin k
// (slot B_ENV, layout make_layout(prev, 2, 0), k-th operator) and the new
// site's spin (the exchange bond of Get_hamiltonian_ndia).  HOP_PRODUCT k is
// sum_sigma c^dag_k c_new = sqrt(2) [c^dag_k x c~_new]^0 of carried c^dag k
// (slot B_ENV, layout make_layout(prev, 1, 1)) and the new site's
// annihilator (the hopping bond), with the fermion sign of `side`: a sys
// site follows its block, an env site precedes it.
// QUAD_PRODUCT k is -[W_k x c~_new]^0 for carried rank-1/2 operator k (slot
// B_ENV, layout make_layout(prev, 1, 1)); with W = P_12 c~_3 it is the
// four-site scalar P_12 Delta_34 of Get_sys_two1.
struct TrackSrc {
  enum What { CARRY, SITE, PRODUCT, SPIN_PRODUCT, HOP_PRODUCT, QUAD_PRODUCT } what;
  int k;
};
// One step of the chain of rank-1/2, N + 1 operators W = P c~_new of
// Get_sys_two1 (sys side): CARRY q embeds old operator q (slot B_SYS,
// layout make_layout(prev, 1, 1)); PAIR_HOLE k multiplies old pair creator k
// (slot B_ENV, layout make_layout(prev, 0, 2)) by the new site's
// annihilator partner.  Output in make_layout(new, 1, 1), q * new_size.
struct HoleSrc {
  enum What { CARRY, PAIR_HOLE } what;
  int k;
};
TrackPlan compile_track_hole(const ModelParams& m, const std::vector<Sector>& prev,
                             const std::vector<KeptSector>& kept, const std::vector<HoleSrc>& srcs);
TrackPlan compile_track_ops(const ModelParams& m, const std::vector<Sector>& prev,
                            const std::vector<KeptSector>& kept, const std::vector<TrackSrc>& srcs, int side = 0);
// rho'_b = (1/sqrt(2S_b+1)) sum over psi blocks and site states of the block
// part of psi (B_IN) for the block sectors `blk` of `side`, to B_OUT in the
// sector-diagonal layout: <O> = Re <O, rho'> for a tracked Hermitian O.  With
// first = B_X the first factor is another vector of the same layout (a
// transition density matrix: Re <O, rho'> = <psi| O |x>).
TaskList compile_block_rho(const SuperBasis& sb, int side, const EnlBasis& eb, const std::vector<Sector>& blk,
                           uint8_t first = B_IN, uint8_t second = B_IN);
// B_OUT = (O_sys (x) 1) psi for a tracked sys-block operator O (slot B_AUX at
// element offset op_off, sector-diagonal layout of `blk`), with O acting as
// O_b / sqrt(2S_b+1) on every (b, s) row block; site = true applies the
// centre-site density n_s instead.  The block form covers every element of
// the output; the site form writes only the rows of site multiplets with
// N > 0, so the caller clears B_OUT first.
TaskList compile_apply_sys(const SuperBasis& sb, const EnlBasis& eb, const std::vector<Sector>& blk, int64_t op_off,
                           bool site);
// B_OUT += n_site psi on the rows (side 0) or columns (side 1) of the centre
// site of `side`: <n_site> = Re <psi, out> with out zeroed beforehand.  With
// what = 1 the site operator is the double occupancy n_up n_down.
TaskList compile_site_density(const SuperBasis& sb, int side, const EnlBasis& eb, int what = 0);

TransPlan compile_wave_trans_rl(const SuperBasis& old_sb, const EnlBasis& old_sys, const EnlBasis& old_env,
                                const std::vector<KeptSector>& sys_trun_of_A,
                                const std::vector<KeptSector>& env_trun,
                                const EnlBasis& new_sys, const EnlBasis& new_env, const SuperBasis& new_sb);

// counter-based start vector (same stream as the oracle's start_vector)
void start_vector(int64_t dim, uint64_t seed, std::vector<cplx>& out);

}  // namespace hh
