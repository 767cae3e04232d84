// Implementation of the drop-in host classes (see xflow_dropin.h).
#include "xflow_dropin.h"

#include <cmath>
#include <cstring>
#include <set>

namespace xflow_dropin {

namespace {
struct Ctx {
  xf_ctx h = nullptr;
  ~Ctx() {
    if (h) xf_shutdown(h);
  }
};
thread_local Ctx tl_ctx;
std::set<std::string> &messages() {
  static std::set<std::string> s;  // keeps thrown strings alive
  return s;
}
}  // namespace

xf_ctx context() {
  if (!tl_ctx.h) check(xf_init(0, &tl_ctx.h), "xf_init");
  return tl_ctx.h;
}

void fail(const std::string &msg) { throw messages().insert(msg).first->c_str(); }

void check(int status, const char *what) {
  if (status == XF_OK) return;
  std::string m = std::string("Error in ") + what + ": " + (tl_ctx.h ? xf_last_error(tl_ctx.h) : "") +
                  " (status " + std::to_string(status) + ")";
  fail(m);
}

}  // namespace xflow_dropin

using xflow_dropin::check;
using xflow_dropin::context;
using xflow_dropin::fail;

// ------------------------------------------------------------------ masks
T3dNumberMask::T3dNumberMask(int N, int L, int M) : n_(N), l_(L), m_(M), data_((size_t)N * L * M, 0) {
  rows_.resize(N);
  planes_.resize(N);
  for (int i = 0; i < N; i++) {
    rows_[i].resize(L);
    for (int j = 0; j < L; j++) rows_[i][j] = data_.data() + ((size_t)i * L + j) * M;
    planes_[i] = rows_[i].data();
  }
}

T3dNumberMask &T3dNumberMask::Copy() const {
  T3dNumberMask *c = new T3dNumberMask(n_, l_, m_);
  c->data_ = data_;
  return *c;
}

// ------------------------------------------------------------- separators
TUniformSeparator::TUniformSeparator(double start, double end, double step) {
  if (start >= end) fail("Error in TSeparator constructor: Start value is equal or more then end value.");
  if (step <= 0) fail("Error in TUniformSeparator constructor: Separation value is equal or less then zero.");
  double resDev = (end - start) / step;
  double iPart = 0;
  std::modf(resDev, &iPart);
  end_index_ = (int)iPart;
  if (std::fabs(resDev - iPart) > 0.00000000001) end_index_ = end_index_ + 1;
  double h = (end - start) / ((double)end_index_);
  dim_.assign(end_index_ + 1, 0.0);
  sep_.assign(end_index_, h);
  dim_[0] = start;
  dim_[end_index_] = end;
  for (int i = 1; i < end_index_; i++) dim_[i] = dim_[i - 1] + h;
}

TRandomSeparator::TRandomSeparator(int count, const double *dims) {
  if (dims[0] >= dims[count - 1]) fail("Error in TSeparator constructor: Start value is equal or more then end value.");
  end_index_ = count - 1;
  dim_.assign(dims, dims + count);
  sep_.assign(end_index_, 0.0);
  sep_[0] = dims[1] - dims[0];
  for (int i = 1; i < end_index_; i++) sep_[i] = dims[i + 1] - dims[i];
}

// ------------------------------------------------------------------ grid
T3dNormalGrid::T3dNormalGrid(TSeparator *s1, TSeparator *s2, TSeparator *s3, const TMask &mask,
                             const T3dVector *const *const *const *normals, bool deleteSeparators)
    : own_(deleteSeparators) {
  s_[0] = s1;
  s_[1] = s2;
  s_[2] = s3;
  const T3dNumberMask *m = dynamic_cast<const T3dNumberMask *>(&mask);
  if (!m) fail("T3dMaskGrid constructor error: Could not cast the mask to T3dNumberMask * type.");
  mask_.reset(&m->Copy());
  const int n = s1->GetEndIndex() + 1, l = s2->GetEndIndex() + 1, mm = s3->GetEndIndex() + 1;
  if (m->GetN() != n || m->GetL() != l || m->GetM() != mm) fail("T3dNormalGrid: mask and separators disagree.");
  if (normals) {
    normals_.assign((size_t)n * l * mm * 3, 0.0);
    has_normal_.assign((size_t)n * l * mm, 0);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < l; j++)
        for (int k = 0; k < mm; k++) {
          const T3dVector *v = normals[i][j][k];
          if (!v) continue;
          size_t c = ((size_t)i * l + j) * mm + k;
          normals_[3 * c] = v->X;
          normals_[3 * c + 1] = v->Y;
          normals_[3 * c + 2] = v->Z;
          has_normal_[c] = 1;
        }
  }
}

const T3dVector *const *const *const *T3dNormalGrid::GetNormals() const {
  if (nplane_.empty()) {
    const int n = s_[0]->GetEndIndex() + 1, l = s_[1]->GetEndIndex() + 1, m = s_[2]->GetEndIndex() + 1;
    const size_t nc = (size_t)n * l * m;
    nvec_.assign(nc, T3dVector{});
    nptr_.assign(nc, nullptr);
    for (size_t c = 0; c < nc && !has_normal_.empty(); c++)
      if (has_normal_[c]) {
        nvec_[c].X = normals_[3 * c];
        nvec_[c].Y = normals_[3 * c + 1];
        nvec_[c].Z = normals_[3 * c + 2];
        nptr_[c] = &nvec_[c];
      }
    nrow_.resize((size_t)n * l);
    for (size_t r = 0; r < nrow_.size(); r++) nrow_[r] = nptr_.data() + r * m;
    nplane_.resize(n);
    for (int i = 0; i < n; i++) nplane_[i] = nrow_.data() + (size_t)i * l;
  }
  return nplane_.data();
}

T3dNormalGrid::~T3dNormalGrid() {
  if (dev_) xf_grid_destroy(context(), dev_);
  if (own_)
    for (auto s : s_) delete s;
}

xf_grid T3dNormalGrid::device_grid() const {
  if (!dev_) {
    const int n = mask_->GetN(), l = mask_->GetL(), m = mask_->GetM();
    std::vector<int32_t> codes(mask_->data(), mask_->data() + (size_t)n * l * m);
    check(xf_grid_create(context(), n, l, m, s_[0]->GetSeparation(), s_[1]->GetSeparation(),
                         s_[2]->GetSeparation(), codes.data(), normals_.empty() ? nullptr : normals_.data(),
                         has_normal_.empty() ? nullptr : has_normal_.data(), &dev_),
          "T3dNormalGrid");
  }
  return dev_;
}

// ------------------------------------------------------------ space model
TRnRelease3dSpaceModel::TRnRelease3dSpaceModel(int dataCount, const TMask *const *masks) {
  if (dataCount != 1) fail("TRnRelease3dSpaceModel: only one data block is supported.");
  mask_ = dynamic_cast<const T3dNumberMask *>(masks[0]);
  if (!mask_) fail("TRnRelease3dSpaceModel constructor error: Could not cast one of the masks to T3dNumberMask * type.");
  const int64_t nc = ncell();
  for (int64_t c = 0; c < nc; c++) {
    const int v = mask_->data()[c];
    if (v != 0 && v != 1)
      fail("TRnRelease3dSpaceModel constructor error:One of the mask element value is incorrect. Must be 0 or 1.");
    end_index_ += v;
  }
}

const T3dNumberMask &TRnRelease3dSpaceModel::GetMask(int id) const {
  if (id != 0) fail("TRnRelease3dSpaceModel::GetMask: unknown data block.");
  return *mask_;
}

int64_t TRnRelease3dSpaceModel::cell_of(int rn) const {
  if (rn_cells_.empty()) {
    const int64_t nc = ncell();
    rn_cells_.reserve(end_index_);
    for (int64_t c = 0; c < nc; c++)
      if (mask_->data()[c] > 0) rn_cells_.push_back(c);
  }
  return rn_cells_[rn - 1];
}

TRnRelease1dSpaceModel::TRnRelease1dSpaceModel(int dataCount, const TMask *const *masks) {
  if (dataCount != 1) fail("TRnRelease1dSpaceModel: only one data block is supported.");
  mask_ = dynamic_cast<const T1dNumberMask *>(masks[0]);
  if (!mask_) fail("TRnRelease1dSpaceModel constructor error: Could not cast one of the masks to T1dNumberMask * type.");
  for (int i = 0; i < mask_->GetN(); i++) {
    const int v = (*mask_)[i];
    if (v != 0 && v != 1)
      fail("TRnRelease1dSpaceModel constructor error:One of the mask element value is incorrect. Must be 0 or 1.");
    if (v != 1) fail("TRnRelease1dSpaceModel: only all-Actual masks are supported (one entry per unknown).");
  }
  end_index_ = mask_->GetN();
}

const T1dNumberMask &TRnRelease1dSpaceModel::GetMask(int id) const {
  if (id != 0) fail("Error in TRnRelease1dSpaceModel::GetMask: The index value is out of range.");
  return *mask_;
}

namespace {
struct ModelEntry {
  TRnSpaceModel *m;
  bool own;
};
std::map<int, ModelEntry> &registry() {
  static std::map<int, ModelEntry> r;
  return r;
}
// a device grid that carries only a space mask (for vector operations when
// no operator has been bound yet): Actual where the mask is 1
struct MaskGrid {
  xf_grid g = nullptr;
};
std::map<const TRnRelease3dSpaceModel *, MaskGrid> &mask_grids() {
  static std::map<const TRnRelease3dSpaceModel *, MaskGrid> r;
  return r;
}
// The device grid used for the vector kernels of a model.  The masked vector
// kernels only read the space mask, so a grid whose point types are the model
// mask (1 Actual, 0 Fictive) serves every model.
const TRnRelease3dSpaceModel *as3d(const TRnSpaceModel &m) { return dynamic_cast<const TRnRelease3dSpaceModel *>(&m); }

xf_grid vec_grid(const TRnRelease3dSpaceModel &m) {
  MaskGrid &mg = mask_grids()[&m];
  if (!mg.g) {
    const T3dNumberMask &mk = m.GetMask(0);
    const int n = mk.GetN(), l = mk.GetL(), mm = mk.GetM();
    std::vector<double> hx(n - 1, 1.0), hy(l - 1, 1.0), hz(mm - 1, 1.0);
    std::vector<int32_t> codes(mk.data(), mk.data() + (size_t)n * l * mm);
    check(xf_grid_create(context(), n, l, mm, hx.data(), hy.data(), hz.data(), codes.data(), nullptr, nullptr,
                         &mg.g),
          "TRnRelease3dSpaceModel");
  }
  return mg.g;
}
}  // namespace

void TRnSpace::AddModel(int name, TRnSpaceModel *model, bool deleteModel) {
  if (!model) fail("TRnSpace::AddModel: null model.");
  if (registry().count(name)) fail("TRnSpace::AddModel: model name already registered.");
  registry()[name] = ModelEntry{model, deleteModel};
}

void TRnSpace::RemoveModel(int name) {
  auto it = registry().find(name);
  if (it == registry().end()) return;
  auto mg = mask_grids().find(as3d(*it->second.m));
  if (mg != mask_grids().end()) {
    xf_grid_destroy(context(), mg->second.g);
    mask_grids().erase(mg);
  }
  if (it->second.own) delete it->second.m;
  registry().erase(it);
}

void TRnSpace::RemoveAllModels() {
  while (!registry().empty()) RemoveModel(registry().begin()->first);
}

TRnSpaceModel *TRnSpace::model(int name) {
  auto it = registry().find(name);
  if (it == registry().end()) fail("TRnSpace: unknown model name.");
  return it->second.m;
}

// ------------------------------------------------------------- vectors
TRnSpace::TRnSpace(int modelName) : model_name_(modelName), model_(model(modelName)) {
  const int64_t nc = model_->ncell();
  host_.assign(nc, 0.0);
  check(xf_malloc(context(), sizeof(double) * nc, (void **)&dev_), "TRnSpace");
  check(xf_memcpy(context(), dev_, host_.data(), sizeof(double) * nc, 0), "TRnSpace");
}

TRnSpace::~TRnSpace() {
  if (dev_) xf_free(context(), dev_);
}

double *TRnSpace::device() const {
  if (host_dirty_) {
    check(xf_memcpy(context(), dev_, host_.data(), sizeof(double) * host_.size(), 0), "TRnSpace upload");
    host_dirty_ = false;
    host_valid_ = true;
  }
  return dev_;
}

void TRnSpace::mark_device_modified() const { host_valid_ = false; }

void TRnSpace::sync_to_host() const {
  if (!host_valid_) {
    check(xf_memcpy(context(), host_.data(), dev_, sizeof(double) * host_.size(), 1), "TRnSpace download");
    host_valid_ = true;
  }
}

static void same_model(const TRnSpace &a, const TRnSpace &b) {
  if (&a.GetModel() != &b.GetModel()) fail("Error in TRnSpace::operator=: models mismatch.");
}

// masked vector algebra: the 3-D kernels over the model's space mask, or the
// Rn kernels over every entry of a 1-D model
static int lincomb_of(const TRnSpaceModel &m, double p1, const double *s1, double p2, const double *s2,
                      double *res) {
  if (const TRnRelease3dSpaceModel *m3 = as3d(m)) return xf_lincomb(context(), vec_grid(*m3), p1, s1, p2, s2, res);
  return xf_rn_lincomb(context(), m.ncell(), p1, s1, p2, s2, res);
}

TRnSpace &TRnSpace::operator|=(const TRnSpace &s) {
  same_model(*this, s);
  const double *src = s.device();
  double *dst = device();
  if (const TRnRelease3dSpaceModel *m3 = as3d(*model_)) check(xf_copy(context(), vec_grid(*m3), src, dst), "operator|=");
  else if (src != dst) check(xf_memcpy(context(), dst, src, sizeof(double) * host_.size(), 2), "operator|=");
  mark_device_modified();
  return *this;
}

TRnSpace &TRnSpace::operator|=(double v) {
  if (v != 0) fail("TRnSpace::operator|=: only 0 can be assigned as a number.");
  check(lincomb_of(*model_, 0, nullptr, 0, nullptr, device()), "operator|=");
  mark_device_modified();
  return *this;
}

TRnSpace &TRnSpace::operator|=(const TLinTerm &t) { return *this |= TLinComb{t, TLinTerm{0, nullptr}}; }

TRnSpace &TRnSpace::operator|=(const TLinComb &t) {
  const double *s1 = t.a.s ? t.a.s->device() : nullptr;
  const double *s2 = t.b.s ? t.b.s->device() : nullptr;
  if (t.a.s) same_model(*this, *t.a.s);
  if (t.b.s) same_model(*this, *t.b.s);
  check(lincomb_of(*model_, t.a.p, s1, t.b.p, s2, device()), "operator|=");
  mark_device_modified();
  return *this;
}

TRnSpace &TRnSpace::operator|=(const TOpApply &t) { return *this |= TOpComb{t, 0.0, nullptr}; }

TRnSpace &TRnSpace::operator|=(const TOpComb &t) {
  if (t.ax.x == this) fail("TRnSpace::operator|=: the operator argument must not be the result vector.");
  const double *x = t.ax.x->device();
  const double *b = t.b ? t.b->device() : nullptr;
  t.ax.A->apply(x, b, t.sign, device());
  mark_device_modified();
  return *this;
}

double TRnSpace::GetAbsMaxElement() {
  int idx;
  return GetAbsMaxElement(idx);
}

double TRnSpace::GetAbsMaxElement(int &elementIndex) {
  double v;
  int32_t idx;
  if (const TRnRelease3dSpaceModel *m3 = as3d(*model_))
    check(xf_absmax(context(), vec_grid(*m3), device(), &v, &idx), "GetAbsMaxElement");
  else
    check(xf_rn_absmax(context(), model_->ncell(), device(), &v, &idx), "GetAbsMaxElement");
  elementIndex = idx;
  return v;
}

TRnSpace &TRnSpace::Assign(const TRnSpace &s) {
  if (model_ != s.model_) fail("Error in TRnSpace::Assign: models mismatch.");
  if (this != &s) {
    check(xf_memcpy(context(), device(), s.device(), sizeof(double) * host_.size(), 2), "Assign");
    mark_device_modified();
  }
  return *this;
}

TRnSpace &TRnSpace::Copy() const {
  TRnSpace &c = CreateInstance();
  check(xf_memcpy(context(), c.dev_, device(), sizeof(double) * host_.size(), 2), "Copy");
  c.mark_device_modified();
  return c;
}

TRnSpace &TRnSpace::CreateInstance() const {
  if (as3d(*model_)) return *new TRnRelease3dSpace(model_name_);
  return *new TRnRelease1dSpace(model_name_);
}

double &TRnSpace::operator[](int rnIndex) const {
  if (rnIndex < 1 || rnIndex > GetEndIndex()) fail("Error in TRnSpace::operator[]: The index is out of range.");
  sync_to_host();
  host_dirty_ = true;  // the caller may write through the reference
  const TRnRelease3dSpaceModel *m3 = as3d(*model_);
  return host_[m3 ? m3->cell_of(rnIndex) : rnIndex - 1];
}

TRnRelease3dSpace::TRnRelease3dSpace(int modelName) : TRnSpace(modelName) {
  const TRnRelease3dSpaceModel *m3 = as3d(*model_);
  if (!m3) fail("TRnRelease3dSpace: the model is not a TRnRelease3dSpaceModel.");
  const T3dNumberMask &mk = m3->GetMask(0);
  rows_.resize(mk.GetN());
  for (int i = 0; i < mk.GetN(); i++) {
    rows_[i].resize(mk.GetL());
    for (int j = 0; j < mk.GetL(); j++) rows_[i][j] = host_.data() + ((size_t)i * mk.GetL() + j) * mk.GetM();
  }
}

double **TRnRelease3dSpace::operator[](int i) {
  sync_to_host();
  host_dirty_ = true;  // the caller may write through the returned rows
  return rows_[i].data();
}

TRnRelease1dSpace::TRnRelease1dSpace(int modelName) : TRnSpace(modelName) {
  if (as3d(*model_)) fail("TRnRelease1dSpace: the model is not a TRnRelease1dSpaceModel.");
}

double &TRnRelease1dSpace::operator[](int i) {
  sync_to_host();
  host_dirty_ = true;
  return host_[i];
}

TLinTerm operator*(double p, const TRnSpace &s) { return TLinTerm{p, &s}; }
TLinComb operator+(const TLinTerm &a, const TLinTerm &b) { return TLinComb{a, b}; }
TLinComb operator-(const TLinTerm &a, const TLinTerm &b) { return TLinComb{a, TLinTerm{-b.p, b.s}}; }
TLinComb operator+(const TRnSpace &a, const TLinTerm &b) { return TLinComb{TLinTerm{1, &a}, b}; }
TLinComb operator-(const TRnSpace &a, const TLinTerm &b) { return TLinComb{TLinTerm{1, &a}, TLinTerm{-b.p, b.s}}; }
TLinComb operator+(const TLinTerm &a, const TRnSpace &b) { return TLinComb{a, TLinTerm{1, &b}}; }
TLinComb operator-(const TLinTerm &a, const TRnSpace &b) { return TLinComb{a, TLinTerm{-1, &b}}; }
TLinComb operator+(const TRnSpace &a, const TRnSpace &b) { return TLinComb{TLinTerm{1, &a}, TLinTerm{1, &b}}; }
TLinComb operator-(const TRnSpace &a, const TRnSpace &b) { return TLinComb{TLinTerm{1, &a}, TLinTerm{-1, &b}}; }
TOpComb operator-(const TOpApply &a, const TRnSpace &b) { return TOpComb{a, -1.0, &b}; }
TOpComb operator+(const TOpApply &a, const TRnSpace &b) { return TOpComb{a, 1.0, &b}; }

double operator,(const TRnSpace &a, const TRnSpace &b) {
  same_model(a, b);
  double v;
  if (const TRnRelease3dSpaceModel *m3 = as3d(a.GetModel()))
    check(xf_dot(context(), vec_grid(*m3), a.device(), b.device(), &v), "operator,");
  else
    check(xf_rn_dot(context(), a.GetModel().ncell(), a.device(), b.device(), &v), "operator,");
  return v;
}

double operator,(const TOpApply &a, const TRnSpace &b) {
  if (a.x == &b) return a.A->apply_dot(b.device());
  // (A x, y) with y != x: materialise A x
  TRnSpace &t = b.CreateInstance();
  t |= a;
  double v = (t, b);
  delete &t;
  return v;
}

// ------------------------------------------------------------- operators
T3dLaplaceOperator::T3dLaplaceOperator(int modelName, int gridsCount, const TNodesGrid *const *grids)
    : model_name_(modelName) {
  if (gridsCount != 1) fail("Error in T3dLaplaceOperator constructor: GridCount<>1.");
  grid_ = dynamic_cast<const T3dNormalGrid *>(grids[0]);
  if (!grid_) fail("Error in T3dLaplaceXOperator constructor: Could not cast grid to T3dNormalGrid* type.");
  // the model mask must be the space mask the grid's point types define
  const TRnRelease3dSpaceModel *m = as3d(*TRnSpace::model(modelName));
  if (!m) fail("T3dLaplaceOperator: the model is not a TRnRelease3dSpaceModel.");
  xf_grid g = grid_->device_grid();
  int64_t nc, nu;
  xf_grid_info(g, &nc, &nu);
  if (nc != m->ncell()) fail("T3dLaplaceOperator: model and grid sizes differ.");
  std::vector<int32_t> sm(nc);
  check(xf_grid_export(context(), g, sm.data(), nullptr, nullptr), "T3dLaplaceOperator");
  const int *md = m->GetMask(0).data();
  for (int64_t c = 0; c < nc; c++)
    if (sm[c] != md[c])
      fail("T3dLaplaceOperator: the model mask differs from the space mask of the grid's point types "
           "(XProblem.cpp:4155-4182).");
}

void T3dLaplaceOperator::ConstructActivityTables() {
  xf_grid g = grid_->device_grid();
  int64_t nc, nu;
  xf_grid_info(g, &nc, &nu);
  act_.assign(nu * 8, 0);
  check(xf_grid_export(context(), g, nullptr, nullptr, act_.data()), "ConstructActivityTables");
}

double T3dLaplaceOperator::operator()(int idData, int i, int j, int k) const {
  if (idData != 0) fail("Error in T3dLaplaceOperator::operator(): Operator is not implemented for multidata index <> 0.");
  if (!last_arg_) fail("T3dLaplaceOperator::operator(): no argument has been given to the operator.");
  const T3dNumberMask &mk = grid_->GetMask();
  xf_grid g = grid_->device_grid();
  int64_t nc, nu;
  xf_grid_info(g, &nc, &nu);
  if (row_cache_src_ != last_arg_->device() || row_cache_.empty()) {
    double *tmp = nullptr;
    check(xf_malloc(context(), sizeof(double) * nc, (void **)&tmp), "operator()");
    check(xf_apply(context(), g, last_arg_->device(), nullptr, 0, tmp), "operator()");
    row_cache_.assign(nc, 0.0);
    check(xf_memcpy(context(), row_cache_.data(), tmp, sizeof(double) * nc, 1), "operator()");
    xf_free(context(), tmp);
    row_cache_src_ = last_arg_->device();
  }
  return row_cache_[((size_t)i * mk.GetL() + j) * mk.GetM() + k];
}

void T3dLaplaceOperator::operator~() const {
  fail("Error in T3dLaplaceOperator::operator~: Operation is not implemented");
}

TRnLinearOperator &T3dLaplaceOperator::Copy() const {
  const TNodesGrid *g[1] = {grid_};
  return *new T3dLaplaceOperator(model_name_, 1, g);
}

void T3dLaplaceOperator::apply(const double *x, const double *b, double sign, double *y) const {
  check(xf_apply(context(), device_grid(), x, b, sign, y), "operator|=");
}

double T3dLaplaceOperator::apply_dot(const double *x) const {
  double v;
  check(xf_apply_dot(context(), device_grid(), x, &v), "operator,");
  return v;
}

void T3dVariableDivergentedOperator::withDensity(TRnRelease3dSpace &density) {
  check(xf_grid_set_density(context(), device_grid(), density.device()), "withDensity");
}

TRnLinearOperator &T3dVariableDivergentedOperator::Copy() const {
  const TNodesGrid *g[1] = {grid_};
  return *new T3dVariableDivergentedOperator(model_name_, 1, g);
}

// ------------------------------------------------------------- packed operator
T1dPackOperator::Handle::~Handle() {
  if (h) xf_pack_destroy(context(), h);
}

T1dPackOperator::T1dPackOperator(int lModelName, const TRnLinearOperator &lA, double **lVector)
    : model_name_(lModelName) {
  const T3dLaplaceOperator *A = dynamic_cast<const T3dLaplaceOperator *>(&lA);
  if (!A) fail("T1dPackOperator: the operator must be a T3dLaplaceOperator or T3dVariableDivergentedOperator.");
  const TRnSpaceModel *m = TRnSpace::model(lModelName);
  if (as3d(*m)) fail("T1dPackOperator: the model must be a TRnRelease1dSpaceModel.");
  const TRnSpaceModel *ma = TRnSpace::model(A->model_name());
  if (ma->GetEndIndex() != m->GetEndIndex())
    fail("Error in T1dPackOperator constructor: Contradiction between size of model and size of constructing matrix.");
  if (lVector && *lVector) fail("Error in T1dPackOperator constructor: Method is not ready for Gausse transformation.");
  pack_ = std::make_shared<Handle>();
  check(xf_pack_create(context(), A->device_grid(), &pack_->h), "T1dPackOperator");
  int64_t nnz, nnzt;
  check(xf_pack_info(pack_->h, &pack_->n, &nnz, &nnzt), "T1dPackOperator");
}

TRnLinearOperator &T1dPackOperator::Copy() const { return *new T1dPackOperator(model_name_, pack_); }

void T1dPackOperator::apply(const double *x, const double *b, double sign, double *y) const {
  if (x == y) fail("T1dPackOperator: the operator argument must not be the result vector.");
  check(xf_pack_apply(context(), pack_->h, transpose_ ? 1 : 0, x, y), "operator|=");
  // A x - b as (A x) + (-1) b: the same bits as the row sum minus b
  if (sign != 0) check(xf_rn_lincomb(context(), pack_->n, 1.0, y, sign, b, y), "operator|=");
}

double T1dPackOperator::apply_dot(const double *x) const {
  double *y = nullptr;
  check(xf_malloc(context(), sizeof(double) * pack_->n, (void **)&y), "operator,");
  int st = xf_pack_apply(context(), pack_->h, transpose_ ? 1 : 0, x, y);
  double v = 0;
  if (st == XF_OK) st = xf_rn_dot(context(), pack_->n, y, x, &v);
  xf_free(context(), y);
  check(st, "operator,");
  return v;
}

double T1dPackOperator::operator()(int idData, int i) const {
  if (idData != 0) fail("Error in T1dPackOperator::operator(): multidata index <> 0.");
  if (!last_arg_) fail("T1dPackOperator::operator(): no argument has been given to the operator.");
  if (i < 0 || i >= pack_->n) fail("T1dPackOperator::operator(): row index out of range.");
  if (row_cache_src_ != last_arg_->device() || row_cache_t_ != transpose_ || row_cache_.empty()) {
    double *tmp = nullptr;
    check(xf_malloc(context(), sizeof(double) * pack_->n, (void **)&tmp), "operator()");
    int st = xf_pack_apply(context(), pack_->h, transpose_ ? 1 : 0, last_arg_->device(), tmp);
    row_cache_.assign(pack_->n, 0.0);
    if (st == XF_OK) st = xf_memcpy(context(), row_cache_.data(), tmp, sizeof(double) * pack_->n, 1);
    xf_free(context(), tmp);
    check(st, "operator()");
    row_cache_src_ = last_arg_->device();
    row_cache_t_ = transpose_;
  }
  return row_cache_[i];
}

// ------------------------------------------------------------------ TSLAE
TSLAE::TSLAE(const TRnLinearOperator &A, TRnSpace &x, TRnSpace &b) : A_(&A), x_(&x), b_(&b) {
  same_model(x, b);
  result_.reset(&x.Copy());  // fResult = &x.Copy(); *fResult |= 0 (XSLAE.cpp:12-13)
  *result_ |= 0;
}

TSLAE::TSLAE(const TRnLinearOperator &A, TRnSpace &x, TRnSpace &u, TRnSpace &v, TRnSpace &w, TRnSpace &b)
    : TSLAE(A, x, b) {
  u_ = &u;
  v_ = &v;
  w_ = &w;
}

TSLAE::~TSLAE() {}

void TSLAE::UpdateA(TRnLinearOperator &A) {
  int name = -1;
  if (const T3dLaplaceOperator *A3 = dynamic_cast<const T3dLaplaceOperator *>(&A)) name = A3->model_name();
  else if (const T1dPackOperator *Ap = dynamic_cast<const T1dPackOperator *>(&A)) name = Ap->model_name();
  if (name != x_->model_name()) fail("Error in TSLAE::UpdateA: the operator acts on another space model.");
  A_ = &A;
}

void TSLAE::solve(int method, double rel_tol, bool project_mean, void *precond) {
  xf_solve_opts o;
  xf_solve_opts_default(method, &o);
  if (rel_tol > 0) o.tol = rel_tol;
  o.project_mean = project_mean ? 1 : 0;
  o.precond = precond;
  int st;
  if (const T3dLaplaceOperator *A3 = dynamic_cast<const T3dLaplaceOperator *>(A_)) {
    st = xf_solve(context(), A3->device_grid(), b_->device(), x_->device(), result_->device(), &o, &stats_);
  } else if (const T1dPackOperator *P = dynamic_cast<const T1dPackOperator *>(A_)) {
    if (method != XF_CG3 && method != XF_BICG && method != XF_MR && method != XF_BICGSTAB)
      fail("TSLAE: the packed operator runs the reference solvers only.");
    st = xf_pack_solve(context(), P->device_pack(), P->transposed() ? 1 : 0, b_->device(), x_->device(),
                       result_->device(), &o, &stats_);
  } else {
    fail("TSLAE: unsupported operator type.");
  }
  result_->mark_device_modified();
  check(st, "TSLAE::Solve");
}

void TSLAE::SolveByScalarRunning(int M, double *u, const double *a, const double *b, const double *c,
                                 const double *f) {
  xf_ctx ctx = context();
  const size_t bytes = sizeof(double) * (M + 1);
  double *d[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  const double *src[5] = {u, a, b, c, f};
  for (int q = 0; q < 5; q++) {
    check(xf_malloc(ctx, bytes, (void **)&d[q]), "SolveByScalarRunning");
    check(xf_memcpy(ctx, d[q], src[q], bytes, 0), "SolveByScalarRunning");
  }
  int st = xf_thomas_batched(ctx, 1, M, d[0], d[1], d[2], d[3], d[4]);
  if (st == XF_OK) st = xf_memcpy(ctx, u, d[0], bytes, 1);
  for (int q = 0; q < 5; q++) xf_free(ctx, d[q]);
  check(st, "SolveByScalarRunning");
}
