// Golden-vector driver for the compiled reference (test infrastructure only).
//
// Links the reference translation units (compiled from /root/reference by
// oracle/Makefile) and drives the hot-path classes on small synthetic inputs,
// so that tests/golden/make_golden.py can store their outputs as fixtures,
// and times the reference CG for bench.py's CPU baseline (cg_timed).  Nothing
// here is shipped; the product never links it.
//
//   ref_harness <command> <in.xfb> <out.xfb>
//
// Commands (each reads named arrays from in.xfb and writes named arrays):
//   op        masks -> space mask, Rn numbering, activity table, A x, A x - b,
//             (A x, x), (x, b), |x|_inf with Rn index
//             (XSpaceModel.cpp:276-411, XOperator.cpp:837-1129, XSpace.cpp:2293-3487)
//   divop     the same through T3dVariableDivergentedOperator (XOperator.cpp:1901-2268)
//   cg|bicg|mr|bicgstab
//             TSLAE solvers (XSLAE.cpp:56-414, 667-767); the iteration count is
//             recovered by counting operator sweeps
//   cg_timed  TSLAE::SolveByConjugateGradients stopped after max_sweeps operator
//             sweeps (the three before the first iteration plus two per
//             iteration); reports the wall time of the sweeps after the first
//             three (`seconds`, steady-state iterations only), `iters` = their
//             iteration count, and `setup_seconds` (masks, Rn model, vectors)
//   thomas    TSLAE::SolveByScalarRunning (XSLAE.cpp:416-445)
//   lod       THydrodynamicsProblem::SolveTransferEquation__ (XProblem.cpp:2784-3475)
//   ib        TImmersedBoundaryProblem::Interpolate / SpreadForce / Index
//             (XProblem.cpp:8550-8791, 8931-9046)
//   pack      T1dPackOperator built from the Laplace operator: packed rows of
//             A and of its transpose (counts, values, 0-based columns) and
//             A x, A^T x on the Rn-ordered x (XOperator.cpp:1135-1571)
//   npy       cnpy::npy_save of a double array (cnpy.h:66-125, 207-236): file bytes
//   vtk_boundary
//             TImmersedBoundaryProblem::OutputBoundary (XProblem.cpp:9543-9635): file bytes
//   pack_bicg|pack_bicgstab|pack_mr
//             TSLAE solvers run on that packed operator over Rn vectors
//             (XSLAE.cpp:159-414, 667-767); sweeps counted per packed row
//   vp_masks  a TVPProblem with the live cylinder-channel statement: the U, V, W,
//             P grids' point types, normals and node coordinates as the
//             reference generates them (XStatement.cpp:7717-7977), the space
//             models' Rn numbering from LocalPrepareVariables
//             (XProblem.cpp:4116-4206), and on P the activity table and A x
//
// File format (.xfb): repeated records
//   u32 name_len, name bytes, u8 dtype ('d' f64 | 'i' i32), u32 ndim,
//   i64 dims[ndim], raw little-endian data.
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <cmath>
#include <map>
#include <string>
#include <vector>
#include <stdexcept>
#include <chrono>

#include "XSLAE.h"
#include "XOperator.h"
#include "XSpace.h"
#include "XSpaceModel.h"
#include "XGrid.h"
#include "XMask.h"
#include "XSeparator.h"
#include "XProblem.h"
#include "XImmersedBoundary.h"
#include "cnpy.h"
#include <unistd.h>

// ----------------------------------------------------------------- xfb I/O
struct Arr {
  char dtype = 'd';
  std::vector<int64_t> dims;
  std::vector<double> d;
  std::vector<int32_t> i;
  int64_t size() const {
    int64_t s = 1;
    for (auto x : dims) s *= x;
    return s;
  }
};
typedef std::map<std::string, Arr> Bag;

static Bag read_xfb(const char* path) {
  Bag bag;
  FILE* fp = fopen(path, "rb");
  if (!fp) throw std::runtime_error(std::string("cannot open ") + path);
  for (;;) {
    uint32_t nl;
    if (fread(&nl, 4, 1, fp) != 1) break;
    std::string name(nl, '\0');
    if (fread(&name[0], 1, nl, fp) != nl) throw std::runtime_error("bad name");
    Arr a;
    uint8_t dt;
    uint32_t nd;
    if (fread(&dt, 1, 1, fp) != 1 || fread(&nd, 4, 1, fp) != 1) throw std::runtime_error("bad hdr");
    a.dtype = (char)dt;
    a.dims.resize(nd);
    if (nd && fread(a.dims.data(), 8, nd, fp) != nd) throw std::runtime_error("bad dims");
    int64_t n = a.size();
    if (a.dtype == 'd') {
      a.d.resize(n);
      if (n && (int64_t)fread(a.d.data(), 8, n, fp) != n) throw std::runtime_error("bad f64 data");
    } else {
      a.i.resize(n);
      if (n && (int64_t)fread(a.i.data(), 4, n, fp) != n) throw std::runtime_error("bad i32 data");
    }
    bag[name] = a;
  }
  fclose(fp);
  return bag;
}

static void write_xfb(const char* path, const Bag& bag) {
  FILE* fp = fopen(path, "wb");
  if (!fp) throw std::runtime_error(std::string("cannot write ") + path);
  for (auto& kv : bag) {
    uint32_t nl = (uint32_t)kv.first.size();
    fwrite(&nl, 4, 1, fp);
    fwrite(kv.first.data(), 1, nl, fp);
    uint8_t dt = (uint8_t)kv.second.dtype;
    fwrite(&dt, 1, 1, fp);
    uint32_t nd = (uint32_t)kv.second.dims.size();
    fwrite(&nd, 4, 1, fp);
    fwrite(kv.second.dims.data(), 8, nd, fp);
    if (dt == 'd') fwrite(kv.second.d.data(), 8, kv.second.d.size(), fp);
    else fwrite(kv.second.i.data(), 4, kv.second.i.size(), fp);
  }
  fclose(fp);
}

static const Arr& need(const Bag& b, const char* name) {
  auto it = b.find(name);
  if (it == b.end()) throw std::runtime_error(std::string("missing input ") + name);
  return it->second;
}
static double scalar(const Bag& b, const char* name) { return need(b, name).d.at(0); }
static int iscalar(const Bag& b, const char* name) { return need(b, name).i.at(0); }

static Arr mk_d(std::vector<int64_t> dims) {
  Arr a; a.dtype = 'd'; a.dims = dims; a.d.assign(a.size(), 0.0); return a;
}
static Arr mk_i(std::vector<int64_t> dims) {
  Arr a; a.dtype = 'i'; a.dims = dims; a.i.assign(a.size(), 0); return a;
}

// ----------------------------------------------------------- grid set-up
// Inputs: x[N+1], y[L+1], z[M+1] node coordinates, codes[(N+1)(L+1)(M+1)],
// optional normals[(N+1)(L+1)(M+1)*3] + has_normal[...].
struct Setup {
  int N = 0, L = 0, M = 0;            // EndIndex of the three separators
  T3dNormalGrid* grid = nullptr;
  T3dNumberMask* space_mask = nullptr;
  TRnRelease3dSpaceModel* model = nullptr;
  int model_name = 0;
};

static int g_model_counter = 700;

static Setup make_setup(const Bag& in, bool all_unknown = false) {
  Setup s;
  const Arr& x = need(in, "x");
  const Arr& y = need(in, "y");
  const Arr& z = need(in, "z");
  const Arr& codes = need(in, "codes");
  s.N = (int)x.size() - 1; s.L = (int)y.size() - 1; s.M = (int)z.size() - 1;
  int n = s.N + 1, l = s.L + 1, m = s.M + 1;
  TSeparator* sx = new TRandomSeparator(n, x.d.data());
  TSeparator* sy = new TRandomSeparator(l, y.d.data());
  TSeparator* sz = new TRandomSeparator(m, z.d.data());
  T3dNumberMask* mask = new T3dNumberMask(n, l, m);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < l; j++)
      for (int k = 0; k < m; k++) (*mask)[i][j][k] = codes.i[((int64_t)i * l + j) * m + k];
  // normals
  T3dVector**** nrm = new T3dVector***[n];
  auto hn = in.find("has_normal");
  auto nv = in.find("normals");
  for (int i = 0; i < n; i++) {
    nrm[i] = new T3dVector**[l];
    for (int j = 0; j < l; j++) {
      nrm[i][j] = new T3dVector*[m];
      for (int k = 0; k < m; k++) {
        nrm[i][j][k] = nullptr;
        int64_t c = ((int64_t)i * l + j) * m + k;
        if (hn != in.end() && hn->second.i[c]) {
          T3dVector* v = new T3dVector;
          v->X = nv->second.d[3 * c + 0];
          v->Y = nv->second.d[3 * c + 1];
          v->Z = nv->second.d[3 * c + 2];
          nrm[i][j][k] = v;
        }
      }
    }
  }
  s.grid = new T3dNormalGrid(sx, sy, sz, *mask, nrm);
  // Space mask: caller-supplied (the derivation rule of XProblem.cpp:4155-4182
  // is applied by the Python side and checked separately) or all ones.
  s.space_mask = new T3dNumberMask(n, l, m);
  auto sm = in.find("space_mask");
  for (int i = 0; i < n; i++)
    for (int j = 0; j < l; j++)
      for (int k = 0; k < m; k++) {
        int64_t c = ((int64_t)i * l + j) * m + k;
        (*s.space_mask)[i][j][k] = all_unknown ? 1 : (sm != in.end() ? sm->second.i[c] : 0);
      }
  TMask** masks = new TMask*[1];
  masks[0] = s.space_mask;
  s.model = new TRnRelease3dSpaceModel(1, masks);
  s.model_name = g_model_counter++;
  TRnSpace::AddModel(s.model_name, s.model, false);
  return s;
}

static TRnRelease3dSpace* make_vec(const Setup& s, const Bag& in, const char* name) {
  TRnRelease3dSpace* v = new TRnRelease3dSpace(s.model_name);
  int n = s.N + 1, l = s.L + 1, m = s.M + 1;
  auto it = in.find(name);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < l; j++)
      for (int k = 0; k < m; k++)
        (*v)[i][j][k] = (it != in.end()) ? it->second.d[((int64_t)i * l + j) * m + k] : 0.0;
  return v;
}

static Arr dump_vec(const Setup& s, TRnSpace& vv) {
  TRnRelease3dSpace& v = dynamic_cast<TRnRelease3dSpace&>(vv);
  int n = s.N + 1, l = s.L + 1, m = s.M + 1;
  Arr a = mk_d({n, l, m});
  for (int i = 0; i < n; i++)
    for (int j = 0; j < l; j++)
      for (int k = 0; k < m; k++) a.d[((int64_t)i * l + j) * m + k] = v[i][j][k];
  return a;
}

// Operator subclasses that (a) count per-cell evaluations, so the number of
// operator sweeps (and hence solver iterations) can be recovered, and (b) make
// operator~ the identity so BiCG can run on the symmetric (uniform-grid,
// Actual/Fictive) masks it is tested on.  Arithmetic is the reference's own.
static long long g_evals = 0;
// cg_timed: wall clock when g_evals reaches g_mark, and a budget of
// evaluations after which the solve is abandoned by a BudgetReached throw
static long long g_mark = -1, g_budget = -1;
static std::chrono::steady_clock::time_point g_t_mark;
struct BudgetReached {};

class CountingLaplace : public T3dLaplaceOperator {
 public:
  CountingLaplace(int model, int gc, const TNodesGrid* const* g)
      : TRnLinearOperator(model), T3dLaplaceOperator(model, gc, g), model_(model), grid_(g[0]) {}
  double operator()(int id, int i, int j, int k) const override {
    if (++g_evals == g_mark) g_t_mark = std::chrono::steady_clock::now();
    if (g_evals == g_budget) throw BudgetReached();
    return T3dLaplaceOperator::operator()(id, i, j, k);
  }
  void operator~() const override {}
  TRnLinearOperator& Copy() const override {
    const TNodesGrid* g[1] = {grid_};
    return *(new CountingLaplace(model_, 1, g));
  }
  int** table() const { return fActivityTable; }

 private:
  int model_;
  const TNodesGrid* grid_;
};

class CountingDivergented : public T3dVariableDivergentedOperator {
 public:
  CountingDivergented(int model, int gc, const TNodesGrid* const* g)
      : TRnLinearOperator(model), T3dVariableDivergentedOperator(model, gc, g) {}
  double operator()(int id, int i, int j, int k) const override {
    ++g_evals;
    return T3dVariableDivergentedOperator::operator()(id, i, j, k);
  }
};

static int count_unknowns(const Setup& s) { return s.model->EndIndex; }

// Exposes the packed rows the reference builds (arithmetic and order are its own).
class PackAccess : public T1dPackOperator {
 public:
  PackAccess(int model, const TRnLinearOperator& A, double** v)
      : TRnLinearOperator(model), T1dPackOperator(model, A, v) {}
  double operator()(int id, int i) const override {
    ++g_evals;
    return T1dPackOperator::operator()(id, i);
  }
  void dump(bool transpose, Arr& counts, Arr& vals, Arr& cols) const {
    const int R = transpose ? fRowsCountT : fRowsCount;
    const int* cnt = transpose ? fRowsElementsCountT : fRowsElementsCount;
    double* const* mv = transpose ? fMatrixPackT : fMatrixPack;
    int* const* mc = transpose ? fColumnsT : fColumns;
    int64_t nnz = 0;
    for (int r = 0; r < R; r++) nnz += cnt[r];
    counts = mk_i({R});
    vals = mk_d({nnz});
    cols = mk_i({nnz});
    int64_t q = 0;
    for (int r = 0; r < R; r++) {
      counts.i[r] = cnt[r];
      for (int e = 0; e < cnt[r]; e++, q++) {
        vals.d[q] = mv[r][e];
        cols.i[q] = mc[r][e];
      }
    }
  }
};



// ----------------------------------------------------------- commands
static void cmd_op(const Bag& in, Bag& out, bool divergented) {
  Setup s = make_setup(in);
  int n = s.N + 1, l = s.L + 1, m = s.M + 1;
  const TNodesGrid* grids[1] = {s.grid};
  TRnLinearOperator* A;
  CountingLaplace* lap = nullptr;
  if (divergented) {
    CountingDivergented* d = new CountingDivergented(s.model_name, 1, grids);
    TRnRelease3dSpace* rho = make_vec(s, in, "rho");
    d->withDensity(*rho);
    A = d;
  } else {
    lap = new CountingLaplace(s.model_name, 1, grids);
    A = lap;
  }
  TRnRelease3dSpace& x = *make_vec(s, in, "xvec");
  TRnRelease3dSpace& b = *make_vec(s, in, "bvec");
  TRnRelease3dSpace& y = *make_vec(s, in, "yinit");
  TRnRelease3dSpace& y2 = *make_vec(s, in, "yinit");
  y |= (*A)(x);
  y2 |= (*A)(x) - b;
  double xax = ((*A)(x), x);
  double xb = (x, b);
  out["Ax"] = dump_vec(s, y);
  out["Ax_minus_b"] = dump_vec(s, y2);
  Arr sc = mk_d({2});
  sc.d[0] = xax; sc.d[1] = xb;
  out["dots"] = sc;
  if (!divergented) {
    // Rn numbering (inverse correspondence), activity table
    int RN = count_unknowns(s);
    Arr rn = mk_i({n, l, m});
    for (int i = 0; i < n; i++)
      for (int j = 0; j < l; j++)
        for (int k = 0; k < m; k++) {
          const TInverseCorrespondence& ic = s.model->GetInverseCorrespondence();
          const T3dInverseCorrespondence& ic3 = dynamic_cast<const T3dInverseCorrespondence&>(ic);
          rn.i[((int64_t)i * l + j) * m + k] = ic3[0][i][j][k];
        }
    out["rn"] = rn;
    lap->ConstructActivityTables();
    int** tab = lap->table();
    Arr at = mk_i({RN, 8});
    for (int r = 0; r < RN; r++) {
      int cnt = tab[r][0];
      at.i[r * 8 + 0] = cnt;
      for (int q = 1; q <= cnt; q++) at.i[r * 8 + q] = tab[r][q];
    }
    out["act"] = at;
    int idx = 0;
    double amax = x.GetAbsMaxElement(idx);
    Arr am = mk_d({1}); am.d[0] = amax; out["absmax"] = am;
    Arr ai = mk_i({1}); ai.i[0] = idx; out["absmax_rn"] = ai;
  }
}

static void cmd_solver(const Bag& in, Bag& out, const std::string& which) {
  Setup s = make_setup(in);
  const TNodesGrid* grids[1] = {s.grid};
  CountingLaplace* A = new CountingLaplace(s.model_name, 1, grids);
  TRnRelease3dSpace* x0 = make_vec(s, in, "x0");
  TRnRelease3dSpace* b = make_vec(s, in, "bvec");
  // TSLAE owns and deletes A, x, b (XSLAE.cpp:21-30); it is never destroyed here.
  TSLAE* slae = new TSLAE(*A, *x0, *b);
  g_evals = 0;
  if (which == "cg") slae->SolveByConjugateGradients();
  else if (which == "bicg") slae->SolveByBiConjugateGradients();
  else if (which == "mr") slae->SolveByMinimalResiduals();
  else if (which == "bicgstab") slae->SolveByBiConjugateGradientsStab();
  else throw std::runtime_error("unknown solver");
  long long RN = count_unknowns(s);
  long long sweeps = RN ? g_evals / RN : 0;
  out["x"] = dump_vec(s, const_cast<TRnSpace&>(slae->GetResult()));
  Arr it = mk_i({2});
  it.i[0] = (int)sweeps;
  it.i[1] = (int)(g_evals % (RN ? RN : 1));
  out["sweeps"] = it;
}

static void cmd_cg_timed(const Bag& in, Bag& out) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  Setup s = make_setup(in);
  const TNodesGrid* grids[1] = {s.grid};
  CountingLaplace* A = new CountingLaplace(s.model_name, 1, grids);
  TRnRelease3dSpace* x0 = make_vec(s, in, "x0");
  TRnRelease3dSpace* b = make_vec(s, in, "bvec");
  TSLAE* slae = new TSLAE(*A, *x0, *b);
  const long long RN = count_unknowns(s);
  const int max_sweeps = iscalar(in, "max_sweeps");
  if (max_sweeps < 5 || (max_sweeps - 3) % 2 != 0) throw std::runtime_error("max_sweeps must be 3 + 2 k, k >= 1");
  const double setup = std::chrono::duration<double>(clk::now() - t0).count();
  g_evals = 0;
  g_mark = 3 * RN;  // the evaluation that starts the fourth sweep
  g_budget = (long long)max_sweeps * RN + 1;  // the first evaluation past the budget
  bool stopped = false;
  try {
    slae->SolveByConjugateGradients();
  } catch (const BudgetReached&) {
    stopped = true;
  }
  const auto t1 = clk::now();
  g_mark = g_budget = -1;
  if (!stopped) throw std::runtime_error("the solve converged inside the sweep budget");
  Arr sec = mk_d({1});
  sec.d[0] = std::chrono::duration<double>(t1 - g_t_mark).count();
  out["seconds"] = sec;
  Arr su = mk_d({1});
  su.d[0] = setup;
  out["setup_seconds"] = su;
  Arr it = mk_i({1});
  it.i[0] = (max_sweeps - 3) / 2;
  out["iters"] = it;
}

static void cmd_pack(const Bag& in, Bag& out) {
  Setup s = make_setup(in);
  const TNodesGrid* grids[1] = {s.grid};
  CountingLaplace* lap = new CountingLaplace(s.model_name, 1, grids);
  const int RN = count_unknowns(s);
  T1dNumberMask* m1 = new T1dNumberMask(RN);
  for (int q = 0; q < RN; q++) (*m1)[q] = TActualPoint;
  TMask** ms = new TMask*[1];
  ms[0] = m1;
  TRnRelease1dSpaceModel* model1 = new TRnRelease1dSpaceModel(1, ms);
  const int name1 = g_model_counter++;
  TRnSpace::AddModel(name1, model1, false);
  double* nullv = nullptr;
  PackAccess* P = new PackAccess(name1, *lap, &nullv);
  P->dump(false, out["counts"], out["vals"], out["cols"]);
  P->dump(true, out["countsT"], out["valsT"], out["colsT"]);
  TRnRelease3dSpace& x3 = *make_vec(s, in, "xvec");
  TRnRelease1dSpace& x1 = *new TRnRelease1dSpace(name1);
  for (int r = 1; r <= RN; r++) x1[r - 1] = ((TRnSpace&)x3)[r];
  TRnRelease1dSpace& y1 = *new TRnRelease1dSpace(name1);
  TRnRelease1dSpace& y2 = *new TRnRelease1dSpace(name1);
  TRnLinearOperator& PA = *P;  // the packed operator through its linear-operator face
  y1 |= PA(x1);
  ~(*P);
  y2 |= PA(x1);
  Arr a = mk_d({RN}), b = mk_d({RN});
  for (int r = 0; r < RN; r++) {
    a.d[r] = y1[r];
    b.d[r] = y2[r];
  }
  out["Ax"] = a;
  out["ATx"] = b;
}

// A TSLAE solver run on the packed operator over Rn vectors (BiCG's A_ = A.Copy()
// then ~A_ selects the transposed rows, XSLAE.cpp:205-206).
static void cmd_pack_solve(const Bag& in, Bag& out, const std::string& which) {
  Setup s = make_setup(in);
  const TNodesGrid* grids[1] = {s.grid};
  CountingLaplace* lap = new CountingLaplace(s.model_name, 1, grids);
  const int RN = count_unknowns(s);
  T1dNumberMask* m1 = new T1dNumberMask(RN);
  for (int q = 0; q < RN; q++) (*m1)[q] = TActualPoint;
  TMask** ms = new TMask*[1];
  ms[0] = m1;
  TRnRelease1dSpaceModel* model1 = new TRnRelease1dSpaceModel(1, ms);
  const int name1 = g_model_counter++;
  TRnSpace::AddModel(name1, model1, false);
  double* nullv = nullptr;
  PackAccess* P = new PackAccess(name1, *lap, &nullv);
  TRnRelease3dSpace& b3 = *make_vec(s, in, "bvec");
  TRnRelease3dSpace& x3 = *make_vec(s, in, "x0");
  TRnRelease1dSpace* b1 = new TRnRelease1dSpace(name1);
  TRnRelease1dSpace* x1 = new TRnRelease1dSpace(name1);
  for (int r = 1; r <= RN; r++) {
    (*b1)[r - 1] = ((TRnSpace&)b3)[r];
    (*x1)[r - 1] = ((TRnSpace&)x3)[r];
  }
  TSLAE* slae = new TSLAE(*P, *x1, *b1);
  g_evals = 0;
  if (which == "pack_bicg") slae->SolveByBiConjugateGradients();
  else if (which == "pack_bicgstab") slae->SolveByBiConjugateGradientsStab();
  else if (which == "pack_mr") slae->SolveByMinimalResiduals();
  else throw std::runtime_error("unknown packed solver");
  const TRnSpace& res = slae->GetResult();
  Arr xr = mk_d({RN});
  for (int r = 1; r <= RN; r++) xr.d[r - 1] = const_cast<TRnSpace&>(res)[r];  // TRnSpace::operator[] is 1-based
  out["x"] = xr;
  Arr it = mk_i({2});
  it.i[0] = (int)(RN ? g_evals / RN : 0);
  it.i[1] = (int)(g_evals % (RN ? RN : 1));
  out["sweeps"] = it;
}

static void cmd_thomas(const Bag& in, Bag& out) {
  const Arr& a = need(in, "a");
  const Arr& b = need(in, "b");
  const Arr& c = need(in, "c");
  const Arr& f = need(in, "f");
  Arr u = need(in, "u");
  int M = (int)u.size() - 1;
  TSLAE::SolveByScalarRunning(M, u.d.data(), a.d.data(), b.d.data(), c.d.data(), f.d.data());
  out["u"] = u;
}

// Exposes the protected transport step.
class LodDriver : public THydrodynamicsProblem {
 public:
  LodDriver(double tau) : THydrodynamicsProblem(1.0, tau, 1.0) {
    fTimeSeparator = new TUniformSeparator(0, 1.0, tau);
  }
  void Solve() override {}
  TNodesGrid* BuildGrid() override { return nullptr; }
  void run(TRnRelease3dSpace& f, TRnRelease3dSpace& f1, TRnRelease3dSpace& u, TRnRelease3dSpace& v,
           TRnRelease3dSpace& w, TRnRelease3dSpace& rp, const TNodesGrid& g, TRnRelease3dSpace& kap, int scheme) {
    SolveTransferEquation__(f, f1, u, v, w, rp, g, kap, scheme);
  }
  double tau() const { return fTimeSeparator->SeparationValue; }
};

static void cmd_lod(const Bag& in, Bag& out) {
  Setup s = make_setup(in);
  LodDriver drv(scalar(in, "tau"));
  TRnRelease3dSpace& f = *make_vec(s, in, "f");
  TRnRelease3dSpace& f1 = *make_vec(s, in, "f_prev");
  TRnRelease3dSpace& u = *make_vec(s, in, "u");
  TRnRelease3dSpace& v = *make_vec(s, in, "v");
  TRnRelease3dSpace& w = *make_vec(s, in, "w");
  TRnRelease3dSpace& rp = *make_vec(s, in, "rp");
  TRnRelease3dSpace& kap = *make_vec(s, in, "kappa");
  drv.run(f, f1, u, v, w, rp, *s.grid, kap, iscalar(in, "scheme"));
  out["f"] = dump_vec(s, f);
  Arr t = mk_d({1}); t.d[0] = drv.tau(); out["tau_used"] = t;
}

class PlainBoundary : public TImmersedBoundary {
 public:
  double GetArea() override { return 0; }
  double GetYCenter() override { return 0; }
  double GetZCenter() override { return 0; }
};

class IbDriver : public TImmersedBoundaryProblem {
 public:
  IbDriver() : TImmersedBoundaryProblem(1.0, 1.0, 1.0, 1.0) {}
  void output_boundary(int step, double dt) { OutputBoundary(step, dt); }
  TMaskGenerator* CreateStatement() override { return nullptr; }
  void set_boundary(TImmersedBoundary* b) { fBoundary = b; }
  void interp(TRnRelease3dSpace& vel, T3dNormalGrid* g, int axis) { Interpolate(vel, g, axis, 0); }
  void spread(TRnRelease3dSpace& f, T3dNormalGrid& gc, T3dNormalGrid* g, int axis) { SpreadForce(f, gc, g, axis, 0); }
  int index(T3dNormalGrid* g, double c, int type, int axis) { return Index(g, c, type, axis, 3, 3); }
};

static void cmd_ib(const Bag& in, Bag& out) {
  Setup s = make_setup(in, /*all_unknown=*/true);
  int axis = iscalar(in, "axis");  // 1, 2, 3 = COORD_X, COORD_Y, COORD_Z
  const Arr& X = need(in, "nodes");  // [n, 3]
  const Arr& F = need(in, "force_n"); // [n]
  int nn = (int)X.dims[0];
  PlainBoundary* bnd = new PlainBoundary();
  bnd->nodesCount = nn;
  bnd->nodes = new TNode*[nn];
  for (int q = 0; q < nn; q++) {
    TNode* t = new TNode();
    t->x = X.d[3 * q]; t->y = X.d[3 * q + 1]; t->z = X.d[3 * q + 2];
    t->type = ELASTIC_NODE;
    t->xForce = t->yForce = t->zForce = F.d[q];
    t->xVel = t->yVel = t->zVel = 0;
    bnd->nodes[q] = t;
  }
  // Allocated and never destroyed: the problem destructors expect a fully
  // initialised problem (grids, statement) that this driver does not build.
  IbDriver& drv = *new IbDriver();
  drv.set_boundary(bnd);
  TRnRelease3dSpace& vel = *make_vec(s, in, "vel");
  TRnRelease3dSpace& force = *make_vec(s, in, "zero");
  drv.interp(vel, s.grid, axis);
  drv.spread(force, *s.grid, s.grid, axis);
  Arr U = mk_d({nn});
  Arr idx = mk_i({nn, 3});
  for (int q = 0; q < nn; q++) {
    TNode* t = bnd->nodes[q];
    U.d[q] = axis == 1 ? t->xVel : axis == 2 ? t->yVel : t->zVel;
    idx.i[3 * q + 0] = drv.index(s.grid, t->x, 1, axis);
    idx.i[3 * q + 1] = drv.index(s.grid, t->y, 2, axis);
    idx.i[3 * q + 2] = drv.index(s.grid, t->z, 3, axis);
  }
  out["U"] = U;
  out["base"] = idx;
  out["force"] = dump_vec(s, force);
}


static Arr file_bytes(const std::string& path) {
  FILE* fp = fopen(path.c_str(), "rb");
  if (!fp) throw std::runtime_error("cannot read " + path);
  std::vector<unsigned char> b;
  int ch;
  while ((ch = fgetc(fp)) != EOF) b.push_back((unsigned char)ch);
  fclose(fp);
  Arr a = mk_i({(int64_t)b.size()});
  for (size_t q = 0; q < b.size(); q++) a.i[q] = b[q];
  return a;
}

// cnpy::npy_save of a double array with the given shape (cnpy.h:66-125)
static void cmd_npy(const Bag& in, Bag& out) {
  const Arr& d = need(in, "data");
  std::vector<unsigned int> shape(d.dims.begin(), d.dims.end());
  char tmpl[] = "/tmp/xf_npy_XXXXXX";
  int fd = mkstemp(tmpl);
  if (fd < 0) throw std::runtime_error("mkstemp");
  close(fd);
  cnpy::npy_save(std::string(tmpl), d.d.data(), shape.data(), (unsigned)shape.size(), "w");
  out["bytes"] = file_bytes(tmpl);
  unlink(tmpl);
}

// TImmersedBoundaryProblem::OutputBoundary (XProblem.cpp:9543-9635): nodes in
// circles of in_circle, each with position, start position, velocity, p, WSS
static void cmd_vtk_boundary(const Bag& in, Bag& out) {
  const Arr& X = need(in, "xyz");
  const Arr& X0 = need(in, "xyz0");
  const Arr& V = need(in, "vel");
  const Arr& P = need(in, "p");
  const Arr& S = need(in, "wss");
  const int circles = iscalar(in, "circles"), per = iscalar(in, "in_circle");
  const int nn = circles * per;
  PlainBoundary* bnd = new PlainBoundary();
  bnd->nodesCount = nn;
  bnd->height_nodes = circles;
  bnd->radius_nodes = per;
  bnd->nodes = new TNode*[nn];
  for (int q = 0; q < nn; q++) {
    TNode* t = new TNode();
    t->x = X.d[3 * q]; t->y = X.d[3 * q + 1]; t->z = X.d[3 * q + 2];
    t->x_0 = X0.d[3 * q]; t->y_0 = X0.d[3 * q + 1]; t->z_0 = X0.d[3 * q + 2];
    t->xVel = V.d[3 * q]; t->yVel = V.d[3 * q + 1]; t->zVel = V.d[3 * q + 2];
    t->p = P.d[q];
    t->ShearStress = S.d[q];
    bnd->nodes[q] = t;
  }
  char dir[] = "/tmp/xf_vtkb_XXXXXX";
  if (!mkdtemp(dir)) throw std::runtime_error("mkdtemp");
  std::string cat = std::string(dir) + "/";
  std::vector<char> catbuf(cat.begin(), cat.end());
  catbuf.push_back('\0');
  TProblem::GlobalCatalog = catbuf.data();
  IbDriver& drv = *new IbDriver();
  drv.set_boundary(bnd);
  const int step = iscalar(in, "step");
  const double dt = scalar(in, "dt");
  drv.output_boundary(step, dt);
  char name[64];
  snprintf(name, sizeof(name), "boundT_%6.4f.vtk", step * dt);
  out["bytes"] = file_bytes(cat + name);
  unlink((cat + name).c_str());
  rmdir(dir);
}

// ------------------------------------------------ vp_masks: the reference's own masks
// A TVPProblem whose statement is the cylinder-channel generator of the live
// problem (TCanalIBMWithElasticBoundary::CreateStatement, XProblem.cpp:10340-10373,
// with its pressure condition, 10360-10366).  Initialize() (XProblem.h:293-306)
// builds the U, V, W, P grids from the generator (XStatement.cpp:7717-7977,
// XProblem.cpp:3989-4108) and runs LocalPrepareVariables (XProblem.cpp:4116-4206),
// which registers each grid's space model; they are read back from the registry.
struct ModelPeek : public TRnRelease3dSpace {
  static TRnSpaceModel* get(int name) { return GetModel(name); }
};

class HarnessVP : public TVPProblem {
 public:
  HarnessVP(double lx, double ly, double lz, int nx, int ny, int nz, double pin, double pout)
      : TVPProblem(1.0, 0.1, 1.0), lx_(lx), ly_(ly), lz_(lz), nx_(nx), ny_(ny), nz_(nz), pin_(pin), pout_(pout) {}
  TMaskGenerator* CreateStatement() override {
    std::function<bool(int, int, int, int, int, int)> condition = [](int i, int j, int k, int L, int M, int K) {
      return (double(j) - double(M) / 2.) * (double(j) - double(M) / 2.) +
                 (double(k) - double(K) / 2.) * (double(k) - double(K) / 2.) >
             (double(M) / 10.1) * (double(K) / 10.1);
    };
    return TSimpleWithCylinderPressureMaskGenerator::CreateInstance(lx_, ly_, lz_, nx_, ny_, nz_, pin_, pout_,
                                                                     condition);
  }
  void run(const Bag& in, Bag& out) {
    Initialize();
    const char* names[4] = {"U", "V", "W", "P"};
    TNodesGrid* grids[4] = {fGridU, fGridV, fGridW, fGridP};
    const int models[4] = {fSpaceModelNameU, fSpaceModelNameV, fSpaceModelNameW, fSpaceModelNameP};
    for (int q = 0; q < 4; q++) {
      T3dNormalGrid& g = dynamic_cast<T3dNormalGrid&>(*grids[q]);
      const int n = g.GetSeparator1().EndIndex + 1, l = g.GetSeparator2().EndIndex + 1,
                m = g.GetSeparator3().EndIndex + 1;
      const std::string nm = names[q];
      const TSeparator* seps[3] = {&g.GetSeparator1(), &g.GetSeparator2(), &g.GetSeparator3()};
      const char* ax[3] = {"_x", "_y", "_z"};
      for (int a = 0; a < 3; a++) {
        const int cnt = seps[a]->EndIndex + 1;
        Arr c = mk_d({cnt});
        const double* dim = seps[a]->GetDimension();
        for (int t = 0; t < cnt; t++) c.d[t] = dim[t];
        out[nm + ax[a]] = c;
      }
      Arr codes = mk_i({n, l, m}), hn = mk_i({n, l, m}), nv = mk_d({n, l, m, 3}), rn = mk_i({n, l, m});
      const T3dNumberMask& mask = g.Mask;
      const T3dVector* const* const* const* nr = g.Normals;
      const TRnRelease3dSpaceModel& model = dynamic_cast<const TRnRelease3dSpaceModel&>(*ModelPeek::get(models[q]));
      const T3dInverseCorrespondence& ic3 =
          dynamic_cast<const T3dInverseCorrespondence&>(model.GetInverseCorrespondence());
      for (int i = 0; i < n; i++)
        for (int j = 0; j < l; j++)
          for (int k = 0; k < m; k++) {
            const int64_t c = ((int64_t)i * l + j) * m + k;
            codes.i[c] = mask[i][j][k];
            rn.i[c] = ic3[0][i][j][k];
            if (nr[i][j][k]) {
              hn.i[c] = 1;
              nv.d[3 * c] = nr[i][j][k]->X;
              nv.d[3 * c + 1] = nr[i][j][k]->Y;
              nv.d[3 * c + 2] = nr[i][j][k]->Z;
            }
          }
      out[nm + "_codes"] = codes;
      out[nm + "_has_normal"] = hn;
      out[nm + "_normals"] = nv;
      out[nm + "_rn"] = rn;
      if (q == 3 && in.count("xvec")) {
        // the pressure operator's rows on this grid and model: activity
        // table and A x for the caller's x (XOperator.cpp:837-1129)
        const TNodesGrid* gp[1] = {&g};
        CountingLaplace* lap = new CountingLaplace(models[q], 1, gp);
        TRnRelease3dSpace& x = *new TRnRelease3dSpace(models[q]);
        TRnRelease3dSpace& y = *new TRnRelease3dSpace(models[q]);
        const Arr& xv = need(in, "xvec");
        if (xv.size() != (int64_t)n * l * m) throw std::runtime_error("vp_masks: xvec must have the P grid's shape");
        for (int i = 0; i < n; i++)
          for (int j = 0; j < l; j++)
            for (int k = 0; k < m; k++) {
              x[i][j][k] = xv.d[((int64_t)i * l + j) * m + k];
              y[i][j][k] = 0;
            }
        TRnLinearOperator* A = lap;
        y |= (*A)(x);
        Arr ax_ = mk_d({n, l, m});
        for (int i = 0; i < n; i++)
          for (int j = 0; j < l; j++)
            for (int k = 0; k < m; k++) ax_.d[((int64_t)i * l + j) * m + k] = y[i][j][k];
        out["P_Ax"] = ax_;
        lap->ConstructActivityTables();
        const int RN = model.EndIndex;
        int** tab = lap->table();
        Arr at = mk_i({RN, 8});
        for (int r = 0; r < RN; r++) {
          const int cnt = tab[r][0];
          at.i[r * 8 + 0] = cnt;
          for (int t = 1; t <= cnt; t++) at.i[r * 8 + t] = tab[r][t];
        }
        out["P_act"] = at;
      }
    }
  }

 private:
  double lx_, ly_, lz_;
  int nx_, ny_, nz_;
  double pin_, pout_;
};

static void cmd_vp_masks(const Bag& in, Bag& out) {
  const Arr& d = need(in, "dims");
  const Arr& len = need(in, "lengths");
  HarnessVP* vp = new HarnessVP(len.d.at(0), len.d.at(1), len.d.at(2), d.i.at(0), d.i.at(1), d.i.at(2),
                                scalar(in, "p_in"), scalar(in, "p_out"));
  vp->run(in, out);
}

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s <command> <in.xfb> <out.xfb>\n", argv[0]);
    return 2;
  }
  // The reference prints progress on stdout; keep stdout for that and
  // report results only through the output file.
  std::string cmd = argv[1];
  try {
    Bag in = read_xfb(argv[2]);
    Bag out;
    if (cmd == "op") cmd_op(in, out, false);
    else if (cmd == "divop") cmd_op(in, out, true);
    else if (cmd == "cg" || cmd == "bicg" || cmd == "mr" || cmd == "bicgstab") cmd_solver(in, out, cmd);
    else if (cmd == "cg_timed") cmd_cg_timed(in, out);
    else if (cmd == "thomas") cmd_thomas(in, out);
    else if (cmd == "lod") cmd_lod(in, out);
    else if (cmd == "ib") cmd_ib(in, out);
    else if (cmd == "pack") cmd_pack(in, out);
    else if (cmd == "pack_bicg" || cmd == "pack_bicgstab" || cmd == "pack_mr") cmd_pack_solve(in, out, cmd);
    else if (cmd == "npy") cmd_npy(in, out);
    else if (cmd == "vtk_boundary") cmd_vtk_boundary(in, out);
    else if (cmd == "vp_masks") cmd_vp_masks(in, out);
    else { fprintf(stderr, "unknown command %s\n", cmd.c_str()); return 2; }
    write_xfb(argv[3], out);
  } catch (const char* e) {
    fprintf(stderr, "reference threw: %s\n", e);
    return 3;
  } catch (const std::exception& e) {
    fprintf(stderr, "harness error: %s\n", e.what());
    return 4;
  }
  return 0;
}
