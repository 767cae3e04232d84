// Drop-in C++ host classes with the reference's call signatures, backed by
// libxflow_hip.so through the C ABI in include/xflow.h.
//
// Reference-style host code (masks, separators, a T3dNormalGrid, an Rn space
// model, TRnRelease3dSpace vectors, T3dLaplaceOperator, TSLAE) compiles
// against this header unchanged at the call sites listed below and runs its
// pressure solve on the GPU:
//   T3dNumberMask(N,L,M), operator[]                    XMask.h:90-131
//   TUniformSeparator / TRandomSeparator                XSeparator.cpp:25-99
//   T3dNormalGrid(sep1, sep2, sep3, mask, normals)      XGrid.h:279-299
//   TRnRelease3dSpaceModel(dataCount, masks)            XSpaceModel.h:499-533
//   TRnSpace::AddModel / RemoveModel / RemoveAllModels  XSpace.h:272-274
//   TRnRelease3dSpace(modelName), operator[]            XSpace.h:595-633
//   x |= a*y + b*z, x |= A(y) - b, (x, y), (A(x), x),
//   GetAbsMaxElement, EndIndex                          XSpace.h:17-221
//   T3dLaplaceOperator(model, gridsCount, grids),
//   operator()(idData,i,j,k), ConstructActivityTables   XOperator.h:292-319
//   T3dVariableDivergentedOperator::withDensity         XOperator.h:410-423
//   T1dNumberMask(N), operator[]                        XMask.h:26-50
//   TRnRelease1dSpaceModel(dataCount, masks)            XSpaceModel.h:288-300
//   TRnRelease1dSpace(modelName), operator[] (0-based),
//   TRnSpace::operator[] (1-based Rn index)             XSpace.h:504-530, XSpace.cpp:911-920
//   T1dPackOperator(model, A, &vector), operator~, Copy,
//   operator()(idData, i)                               XOperator.h:322-370
//   Property accessors: mask.N/L/M/Mask, sep.EndIndex/CountNodes/StartValue/
//   EndValue/Dimension/Separation, grid.Separator1-3/Mask/Normals,
//   model.EndIndex, x.EndIndex, slae.Result           property.h, XMask.h:102-120,
//                                                       XSeparator.h:18-54, XGrid.h,
//                                                       XSpaceModel.h:50-54, XSpace.h:288-292
//   TRnSpace::Assign                                    XSpace.cpp:813-819
//   TSLAE(A, x, b) and TSLAE(A, x, u, v, w, b), Result / GetResult,
//   UpdateP/UpdateRP (Assign), UpdateA,
//   SolveByConjugateGradients, SolveByBiConjugateGradients,
//   SolveByMinimalResiduals, SolveByBiConjugateGradientsStab,
//   static SolveByScalarRunning                         XSLAE.h:9-48
// Differences, by design: TSLAE does not own (delete) its arguments (the
// reference's destructor deletes them, XSLAE.cpp:21-30, which double-frees at
// XProblem.cpp:767-768); the space mask of a model used with an operator must
// be the one XProblem.cpp:4155-4182 derives from the grid's point types;
// solvers stop with an error instead of looping forever on NaN or stagnation;
// a 1-D model's mask must be all Actual (the reference's only use of one,
// XProblem.cpp:229-235, 712-718); BiCG on a T3dLaplaceOperator takes A^T = A
// (the reference's operator~ throws there, XOperator.h:302).
// Errors are thrown as `const char*`, the reference's convention.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/xflow.h"

const int TFictivePoint = 0;
const int TActualPoint = 1;
const int TBorderPoint = 2;
const int TDefinedBorderPoint = 3;
const int TPreDefinedBorderPoint = 4;
const int TNormalBorderPoint = 5;
const int TPreNormalBorderPoint = 6;
const int TEquationBorderPoint = 7;
const int TPreEquationBorderPoint = 8;

namespace xflow_dropin {
xf_ctx context();                                // one per host thread, device 0
[[noreturn]] void fail(const std::string &msg);  // throws a stable const char*
void check(int status, const char *what);
}  // namespace xflow_dropin

// Read-only member that reads like a field and calls its owner's getter, as
// the reference's Property<C, T, getter, setter> does (property.h:4-41) for
// `mask.N`, `x.EndIndex`, `slae.Result`.  Its setters all throw "Not
// implemented" in the reference; here assignment does not compile.  Owners
// are not copyable (the reference's copy constructors are private or throw),
// so the owner pointer cannot go stale.
template <class C, class T, T (C::*Get)() const>
class Property {
 public:
  explicit Property(const C *owner) : owner_(owner) {}
  Property(const Property &) = delete;
  Property &operator=(const Property &) = delete;
  operator T() const { return (owner_->*Get)(); }
  T operator()() const { return (owner_->*Get)(); }

 private:
  const C *owner_;
};

// ------------------------------------------------------------------ masks
class TMask {
 public:
  virtual ~TMask() {}
};

class T3dNumberMask : public TMask {
 public:
  T3dNumberMask(int N, int L, int M);
  T3dNumberMask(const T3dNumberMask &) = delete;
  T3dNumberMask &operator=(const T3dNumberMask &) = delete;
  T3dNumberMask &Copy() const;
  int *const *operator[](int i) const { return rows_[i].data(); }
  int GetN() const { return n_; }
  int GetL() const { return l_; }
  int GetM() const { return m_; }
  int ***GetMask() const { return const_cast<int ***>(planes_.data()); }
  // XMask.h:102-120
  Property<T3dNumberMask, int, &T3dNumberMask::GetN> N{this};
  Property<T3dNumberMask, int, &T3dNumberMask::GetL> L{this};
  Property<T3dNumberMask, int, &T3dNumberMask::GetM> M{this};
  Property<T3dNumberMask, int ***, &T3dNumberMask::GetMask> Mask{this};
  const int *data() const { return data_.data(); }

 private:
  int n_, l_, m_;
  std::vector<int> data_;
  std::vector<std::vector<int *>> rows_;
  std::vector<int **> planes_;
};

class T1dNumberMask : public TMask {
 public:
  explicit T1dNumberMask(int N) : data_((size_t)(N > 0 ? N : 0), 0) {}
  T1dNumberMask(const T1dNumberMask &) = delete;
  T1dNumberMask &operator=(const T1dNumberMask &) = delete;
  T1dNumberMask &Copy() const {
    T1dNumberMask *c = new T1dNumberMask(GetN());
    c->data_ = data_;
    return *c;
  }
  int &operator[](int i) { return data_[i]; }
  int operator[](int i) const { return data_[i]; }
  int GetN() const { return (int)data_.size(); }
  Property<T1dNumberMask, int, &T1dNumberMask::GetN> N{this};  // XMask.h:30-35
  const int *data() const { return data_.data(); }

 private:
  std::vector<int> data_;
};

// ------------------------------------------------------------- separators
class TSeparator {
 public:
  TSeparator() = default;
  TSeparator(const TSeparator &) = delete;
  TSeparator &operator=(const TSeparator &) = delete;
  virtual ~TSeparator() {}
  int GetEndIndex() const { return end_index_; }
  int GetCountNodes() const { return end_index_ + 1; }
  double GetStartValue() const { return dim_.empty() ? 0.0 : dim_.front(); }
  double GetEndValue() const { return dim_.empty() ? 0.0 : dim_.back(); }
  const double *GetSeparation() const { return sep_.data(); }
  const double *GetDimension() const { return dim_.data(); }
  // XSeparator.h:18-54
  Property<TSeparator, double, &TSeparator::GetStartValue> StartValue{this};
  Property<TSeparator, double, &TSeparator::GetEndValue> EndValue{this};
  Property<TSeparator, int, &TSeparator::GetEndIndex> EndIndex{this};
  Property<TSeparator, int, &TSeparator::GetCountNodes> CountNodes{this};
  Property<TSeparator, const double *, &TSeparator::GetDimension> Dimension{this};
  Property<TSeparator, const double *, &TSeparator::GetSeparation> Separation{this};

 protected:
  int end_index_ = 0;
  std::vector<double> sep_, dim_;
};

// XSeparator.cpp:25-64: node count rounded up with a 1e-11 slack, step
// recomputed, coordinates by cumulative addition.
class TUniformSeparator : public TSeparator {
 public:
  TUniformSeparator(double start, double end, double step);
};

// XSeparator.cpp:81-99
class TRandomSeparator : public TSeparator {
 public:
  TRandomSeparator(int count, const double *dims);
};

struct T2dVector {
  double X, Y;
};
struct T3dVector : public T2dVector {
  double Z;
};

class TNodesGrid {
 public:
  virtual ~TNodesGrid() {}
};

class T3dNormalGrid : public TNodesGrid {
 public:
  T3dNormalGrid(TSeparator *s1, TSeparator *s2, TSeparator *s3, const TMask &mask,
                const T3dVector *const *const *const *normals, bool deleteSeparators = true);
  T3dNormalGrid(const T3dNormalGrid &) = delete;
  T3dNormalGrid &operator=(const T3dNormalGrid &) = delete;
  ~T3dNormalGrid();
  const TSeparator &GetSeparator1() const { return *s_[0]; }
  const TSeparator &GetSeparator2() const { return *s_[1]; }
  const TSeparator &GetSeparator3() const { return *s_[2]; }
  const T3dNumberMask &GetMask() const { return *mask_; }
  // per-node outward normal or NULL (XGrid.h:279-299)
  const T3dVector *const *const *const *GetNormals() const;
  // XGrid.h:45-47, 71-73, 96-98, 204, 282-299
  Property<T3dNormalGrid, const TSeparator &, &T3dNormalGrid::GetSeparator1> Separator1{this};
  Property<T3dNormalGrid, const TSeparator &, &T3dNormalGrid::GetSeparator2> Separator2{this};
  Property<T3dNormalGrid, const TSeparator &, &T3dNormalGrid::GetSeparator3> Separator3{this};
  Property<T3dNormalGrid, const T3dNumberMask &, &T3dNormalGrid::GetMask> Mask{this};
  Property<T3dNormalGrid, const T3dVector *const *const *const *, &T3dNormalGrid::GetNormals> Normals{this};
  // device grid for a given space mask (built on first use)
  xf_grid device_grid() const;

 private:
  TSeparator *s_[3];
  bool own_;
  std::unique_ptr<T3dNumberMask> mask_;
  std::vector<double> normals_;
  std::vector<int32_t> has_normal_;
  mutable xf_grid dev_ = nullptr;
  // GetNormals' pointer view of normals_ (built on first use)
  mutable std::vector<T3dVector> nvec_;
  mutable std::vector<const T3dVector *> nptr_;
  mutable std::vector<const T3dVector *const *> nrow_;
  mutable std::vector<const T3dVector *const *const *> nplane_;
};

// ------------------------------------------------------------ space model
class TRnSpaceModel {
 public:
  TRnSpaceModel() = default;
  TRnSpaceModel(const TRnSpaceModel &) = delete;
  TRnSpaceModel &operator=(const TRnSpaceModel &) = delete;
  virtual ~TRnSpaceModel() {}
  int GetEndIndex() const { return end_index_; }
  Property<TRnSpaceModel, int, &TRnSpaceModel::GetEndIndex> EndIndex{this};  // XSpaceModel.h:50-54
  virtual int64_t ncell() const = 0;  // entries of one vector of the model

 protected:
  int end_index_ = 0;
};

class TRnRelease3dSpaceModel : public TRnSpaceModel {
 public:
  TRnRelease3dSpaceModel(int dataCount, const TMask *const *masks);
  const T3dNumberMask &GetMask(int id) const;
  int64_t ncell() const override { return (int64_t)mask_->GetN() * mask_->GetL() * mask_->GetM(); }
  // array cell of the unknown with 1-based Rn index rn (i->j->k numbering)
  int64_t cell_of(int rn) const;

 private:
  const T3dNumberMask *mask_;
  mutable std::vector<int64_t> rn_cells_;
};

// Rn vectors of one data block over an all-Actual T1dNumberMask
class TRnRelease1dSpaceModel : public TRnSpaceModel {
 public:
  TRnRelease1dSpaceModel(int dataCount, const TMask *const *masks);
  const T1dNumberMask &GetMask(int id) const;
  int64_t ncell() const override { return mask_->GetN(); }

 private:
  const T1dNumberMask *mask_;
};

// ------------------------------------------------------------- vectors
class TRnLinearOperator;
class TRnSpace;
class TRnRelease3dSpace;

struct TLinTerm {  // p * s
  double p;
  const TRnSpace *s;
};
struct TLinComb {  // p1*s1 + p2*s2
  TLinTerm a, b;
};
struct TOpApply {  // A(x)
  const TRnLinearOperator *A;
  const TRnSpace *x;
};
struct TOpComb {  // A(x) + sign*b
  TOpApply ax;
  double sign;
  const TRnSpace *b;
};

class TRnSpace {
 public:
  virtual ~TRnSpace();
  static void AddModel(int name, TRnSpaceModel *model, bool deleteModel = false);
  static void RemoveModel(int name);
  static void RemoveAllModels();
  static TRnSpaceModel *model(int name);

  TRnSpace(const TRnSpace &) = delete;
  TRnSpace &operator=(const TRnSpace &) = delete;
  int GetEndIndex() const { return model_->GetEndIndex(); }
  Property<TRnSpace, int, &TRnSpace::GetEndIndex> EndIndex{this};  // XSpace.h:288-292
  TRnSpace &operator|=(const TRnSpace &s);   // masked copy
  // every entry, unknowns or not; the models must match (XSpace.cpp:813-819, 33-41)
  TRnSpace &Assign(const TRnSpace &s);
  TRnSpace &operator|=(double zero);         // only 0 is supported (XSpace.h:92-102)
  TRnSpace &operator|=(const TLinTerm &t);
  TRnSpace &operator|=(const TLinComb &t);
  TRnSpace &operator|=(const TOpApply &t);
  TRnSpace &operator|=(const TOpComb &t);
  double GetAbsMaxElement();
  double GetAbsMaxElement(int &elementIndex);
  TRnSpace &Copy() const;            // full copy, caller owns
  TRnSpace &CreateInstance() const;  // zero vector of the same model, caller owns
  // host element by 1-based Rn index (XSpace.cpp:911-920)
  double &operator[](int rnIndex) const;

  // device access
  double *device() const;            // uploads pending host writes
  const TRnSpaceModel &GetModel() const { return *model_; }
  int model_name() const { return model_name_; }
  void mark_device_modified() const;

 protected:
  explicit TRnSpace(int modelName);
  int model_name_;
  TRnSpaceModel *model_;
  mutable double *dev_ = nullptr;
  mutable std::vector<double> host_;
  mutable bool host_valid_ = true;   // host copy equals the device copy
  mutable bool host_dirty_ = false;  // host copy was handed out for writing
  void sync_to_host() const;
  friend class TRnRelease3dSpace;
  friend class TRnRelease1dSpace;
};

class TRnRelease3dSpace : public TRnSpace {
 public:
  explicit TRnRelease3dSpace(int modelName);
  double **operator[](int i);  // host view; device copy refreshed on next device use

 private:
  std::vector<std::vector<double *>> rows_;
};

class TRnRelease1dSpace : public TRnSpace {
 public:
  explicit TRnRelease1dSpace(int modelName);
  double &operator[](int i);  // 0-based host view; device copy refreshed on next device use
};

TLinTerm operator*(double p, const TRnSpace &s);
TLinComb operator+(const TLinTerm &a, const TLinTerm &b);
TLinComb operator-(const TLinTerm &a, const TLinTerm &b);
TLinComb operator+(const TRnSpace &a, const TLinTerm &b);
TLinComb operator-(const TRnSpace &a, const TLinTerm &b);
TLinComb operator+(const TLinTerm &a, const TRnSpace &b);
TLinComb operator-(const TLinTerm &a, const TRnSpace &b);
TLinComb operator+(const TRnSpace &a, const TRnSpace &b);
TLinComb operator-(const TRnSpace &a, const TRnSpace &b);
TOpComb operator-(const TOpApply &a, const TRnSpace &b);
TOpComb operator+(const TOpApply &a, const TRnSpace &b);
double operator,(const TRnSpace &a, const TRnSpace &b);
double operator,(const TOpApply &a, const TRnSpace &b);

// ------------------------------------------------------------- operators
class TRnLinearOperator {
 public:
  virtual ~TRnLinearOperator() {}
  virtual TOpApply operator()(const TRnSpace &x) const = 0;
  virtual void operator~() const = 0;                  // XOperator.h:72
  virtual TRnLinearOperator &Copy() const = 0;         // XOperator.h:170
  // device products on the unknowns: y = A x + sign * b (b may be null when
  // sign is 0), and (A x, x)
  virtual void apply(const double *x, const double *b, double sign, double *y) const = 0;
  virtual double apply_dot(const double *x) const = 0;
};

class T3dLaplaceOperator : public TRnLinearOperator {
 public:
  T3dLaplaceOperator(int modelName, int gridsCount, const TNodesGrid *const *grids);
  TOpApply operator()(const TRnSpace &x) const override {
    last_arg_ = &x;
    return TOpApply{this, &x};
  }
  void operator~() const override;  // throws, as XOperator.h:302
  TRnLinearOperator &Copy() const override;
  void apply(const double *x, const double *b, double sign, double *y) const override;
  double apply_dot(const double *x) const override;
  // one row of A at (i, j, k) for the vector last passed to operator()(x)
  // (the reference's per-cell virtual call, XOperator.cpp:938-1129)
  double operator()(int idData, int i, int j, int k) const;
  void ConstructActivityTables();
  const std::vector<int32_t> &ActivityTable() const { return act_; }
  xf_grid device_grid() const { return grid_->device_grid(); }
  int model_name() const { return model_name_; }

 protected:
  int model_name_;
  const T3dNormalGrid *grid_;
  std::vector<int32_t> act_;
  mutable const TRnSpace *last_arg_ = nullptr;
  mutable std::vector<double> row_cache_;
  mutable const double *row_cache_src_ = nullptr;
};

class T3dVariableDivergentedOperator : public T3dLaplaceOperator {
 public:
  T3dVariableDivergentedOperator(int modelName, int gridsCount, const TNodesGrid *const *grids)
      : T3dLaplaceOperator(modelName, gridsCount, grids) {}
  void withDensity(TRnRelease3dSpace &density);
  TRnLinearOperator &Copy() const override;  // shares the device grid and its density
};

// Packed rows of a 3-D operator and of its transpose over the Rn unknowns
// (XOperator.cpp:1135-1571), held on the device (xf_pack).  The model is a
// 1-D model with one entry per unknown; *lVector must be NULL (the reference
// throws otherwise, XOperator.cpp:1374-1376).  Copy() shares the rows and
// starts untransposed; operator~ toggles which matrix products use.
class T1dPackOperator : public TRnLinearOperator {
 public:
  T1dPackOperator(int lModelName, const TRnLinearOperator &lA, double **lVector);
  TOpApply operator()(const TRnSpace &x) const override {
    last_arg_ = &x;
    return TOpApply{this, &x};
  }
  // entry i (0-based) of the product with the vector last passed to operator()(x)
  double operator()(int idData, int i) const;
  void operator~() const override { transpose_ = !transpose_; }
  TRnLinearOperator &Copy() const override;
  void apply(const double *x, const double *b, double sign, double *y) const override;
  double apply_dot(const double *x) const override;
  xf_pack device_pack() const { return pack_->h; }
  bool transposed() const { return transpose_; }
  int model_name() const { return model_name_; }

 private:
  struct Handle {
    xf_pack h = nullptr;
    int64_t n = 0;
    ~Handle();
  };
  T1dPackOperator(int modelName, std::shared_ptr<Handle> pack) : model_name_(modelName), pack_(std::move(pack)) {}
  int model_name_;
  std::shared_ptr<Handle> pack_;
  mutable bool transpose_ = false;
  mutable const TRnSpace *last_arg_ = nullptr;
  mutable std::vector<double> row_cache_;
  mutable const double *row_cache_src_ = nullptr;
  mutable bool row_cache_t_ = false;
};

// ------------------------------------------------------------------ TSLAE
class TSLAE {
 public:
  TSLAE(const TRnLinearOperator &A, TRnSpace &x, TRnSpace &b);
  // XSLAE.h:13: u, v, w are held and unused, as in the reference (no solver reads them)
  TSLAE(const TRnLinearOperator &A, TRnSpace &x, TRnSpace &u, TRnSpace &v, TRnSpace &w, TRnSpace &b);
  TSLAE(const TSLAE &) = delete;
  TSLAE &operator=(const TSLAE &) = delete;
  virtual ~TSLAE();
  const TRnSpace &GetResult() const { return *result_; }
  Property<TSLAE, const TRnSpace &, &TSLAE::GetResult> Result{this};  // XSLAE.h:16-20
  static void SolveByScalarRunning(int M, double *u, const double *a, const double *b, const double *c,
                                   const double *f);
  // x and b take every entry of p and rp (TRnSpace::Assign, XSLAE.cpp:40-50)
  void UpdateP(TRnSpace &p) { x_->Assign(p); }
  void UpdateRP(TRnSpace &rp) { b_->Assign(rp); }
  // the operator of later solves (the reference declares it, XSLAE.h:30, and
  // throws "too lazy to make this", XSLAE.cpp:52-55; here it replaces A, which
  // must act on the same model as x)
  void UpdateA(TRnLinearOperator &A);
  void SolveByConjugateGradients() { solve(XF_CG3); }
  void SolveByBiConjugateGradients() { solve(XF_BICG); }
  void SolveByMinimalResiduals() { solve(XF_MR); }
  void SolveByBiConjugateGradientsStab() { solve(XF_BICGSTAB); }
  // extensions: the GPU throughput, mixed-precision and learned-preconditioner paths
  void SolveByFastConjugateGradients(double rel_tol = 1e-8, bool project_mean = true) {
    solve(XF_CG, rel_tol, project_mean);
  }
  void SolveByMixedPrecision(double rel_tol = 1e-8, bool project_mean = true) {
    solve(XF_CG_MP, rel_tol, project_mean);
  }
  // flexible CG in FP64 refinement with the V-cycle preconditioner; `net` (an
  // xf_unet, e.g. with saved parameters loaded) is the solve of the cycle cut
  // at 64^3, or nullptr for the full-depth classical cycle (the faster one)
  void SolveByLearnedPreconditioner(xf_unet net, double rel_tol = 1e-8, bool project_mean = true) {
    solve(XF_FCG_MGNN, rel_tol, project_mean, net);
  }
  // the reference's BiCGStab (recurrence, max-norm stop at its 1e-5 absolute
  // tolerance, operator rows) in five fused passes per iteration
  void SolveByFastBiConjugateGradientsStab() { solve(XF_BICGSTAB_T); }
  const xf_solve_stats &LastStats() const { return stats_; }

 private:
  void solve(int method, double rel_tol = 0, bool project_mean = false, void *precond = nullptr);
  const TRnLinearOperator *A_;
  TRnSpace *x_, *b_;
  TRnSpace *u_ = nullptr, *v_ = nullptr, *w_ = nullptr;
  std::unique_ptr<TRnSpace> result_;
  xf_solve_stats stats_{};
};
