// viennals::Advect<T,D> and friends on the MI355X path: a header-only C++
// facade over the C ABI of libviennals_gpu.so (include/viennals_gpu.h) with
// the names, setters and error behaviour of the reference
// (include/viennals/lsAdvect.hpp:54-1044, lsDomain.hpp, lsMakeGeometry.hpp,
// lsExpand.hpp, lsReduce.hpp, lsPrune.hpp, lsBooleanOperation.hpp).
//
// Differences a caller sees:
//  * T must be double (the GPU path computes in FP64); D is 2 or 3.
//  * Velocities: a BuiltinVelocityField runs as a GPU functor; any other
//    subclass of VelocityField<T> is evaluated on the host through its
//    virtual methods for every active point of every stage (host callbacks
//    of the library).
//  * Errors are logged to stderr and the call returns, as the reference does
//    with its Logger (lsAdvect.hpp:1019-1027).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <functional>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../viennals_gpu.h"

namespace viennals {

template <class T> class SmartPointer : public std::shared_ptr<T> {
public:
  SmartPointer() = default;
  SmartPointer(std::nullptr_t) {}
  SmartPointer(const std::shared_ptr<T> &p) : std::shared_ptr<T>(p) {}
  template <class U> SmartPointer(const SmartPointer<U> &p) : std::shared_ptr<T>(p) {}
  template <class... Args> static SmartPointer New(Args &&...args) {
    return SmartPointer(std::make_shared<T>(std::forward<Args>(args)...));
  }
};

template <class T> using Vec3D = std::array<T, 3>;

namespace detail {
inline bool ok(int code, const char *what) {
  if (code != 0) {
    std::fprintf(stderr, "[viennals_gpu] %s: %s\n", what, lsg_last_error());
    return false;
  }
  return true;
}
}  // namespace detail

enum class BoundaryConditionEnum : unsigned {
  REFLECTIVE_BOUNDARY = 0,
  INFINITE_BOUNDARY = 1,
  PERIODIC_BOUNDARY = 2,
  POS_INFINITE_BOUNDARY = 3,
  NEG_INFINITE_BOUNDARY = 4
};

enum class SpatialSchemeEnum : unsigned {
  ENGQUIST_OSHER_1ST_ORDER = 0,
  ENGQUIST_OSHER_2ND_ORDER = 1,
  LAX_FRIEDRICHS_1ST_ORDER = 2,
  LAX_FRIEDRICHS_2ND_ORDER = 3,
  LOCAL_LAX_FRIEDRICHS_ANALYTICAL_1ST_ORDER = 4,
  LOCAL_LOCAL_LAX_FRIEDRICHS_1ST_ORDER = 5,
  LOCAL_LOCAL_LAX_FRIEDRICHS_2ND_ORDER = 6,
  LOCAL_LAX_FRIEDRICHS_1ST_ORDER = 7,
  LOCAL_LAX_FRIEDRICHS_2ND_ORDER = 8,
  STENCIL_LOCAL_LAX_FRIEDRICHS_1ST_ORDER = 9,
  WENO_3RD_ORDER = 10,
  WENO_5TH_ORDER = 11
};
using IntegrationSchemeEnum = SpatialSchemeEnum;

enum class TemporalSchemeEnum : unsigned { FORWARD_EULER = 0, RUNGE_KUTTA_2ND_ORDER = 1, RUNGE_KUTTA_3RD_ORDER = 2 };

enum struct BooleanOperationEnum : unsigned { INTERSECT = 0, UNION = 1, RELATIVE_COMPLEMENT = 2 };

// ------------------------------------------------------------- Domain ---

// Point data of a Domain (ViennaCore PointData, as used by lsDomain.hpp:44):
// named scalar arrays (one value per defined point, point-id order) and
// vector arrays (three values per point), resident on the device and carried
// through every band operation and Advect.  The getters return copies.
template <class T> class PointData {
  lsg_domain *h_;

public:
  using ScalarDataType = std::vector<T>;
  using VectorDataType = std::vector<std::array<T, 3>>;
  explicit PointData(lsg_domain *h) : h_(h) {}

  void insertNextScalarData(const ScalarDataType &s, const std::string &label = "Scalars") { put(s.data(), s.size(), 1, label, 0); }
  void insertNextVectorData(const VectorDataType &v, const std::string &label = "Vectors") { put(v.data()->data(), v.size(), 3, label, 0); }
  void insertReplaceScalarData(const ScalarDataType &s, const std::string &label = "Scalars") { put(s.data(), s.size(), 1, label, 1); }
  void insertReplaceVectorData(const VectorDataType &v, const std::string &label = "Vectors") { put(v.data()->data(), v.size(), 3, label, 1); }
  unsigned getScalarDataSize() const { return (unsigned)lsg_domain_point_data_count(h_, 1); }
  unsigned getVectorDataSize() const { return (unsigned)lsg_domain_point_data_count(h_, 3); }
  std::string getScalarDataLabel(int i) const { return label(1, i); }
  std::string getVectorDataLabel(int i) const { return label(3, i); }
  std::shared_ptr<ScalarDataType> getScalarData(int i) const {
    auto out = std::make_shared<ScalarDataType>();
    if (!fetch(1, i, out->data(), 0, out.get())) return nullptr;
    return out;
  }
  std::shared_ptr<ScalarDataType> getScalarData(const std::string &l, bool noWarning = false) const {
    const int i = find(1, l, noWarning);
    return i < 0 ? nullptr : getScalarData(i);
  }
  std::shared_ptr<VectorDataType> getVectorData(int i) const {
    auto out = std::make_shared<VectorDataType>();
    std::vector<T> flat;
    if (!fetch(3, i, nullptr, 0, &flat)) return nullptr;
    out->resize(flat.size() / 3);
    for (size_t k = 0; k < out->size(); ++k) (*out)[k] = {flat[3 * k], flat[3 * k + 1], flat[3 * k + 2]};
    return out;
  }
  std::shared_ptr<VectorDataType> getVectorData(const std::string &l, bool noWarning = false) const {
    const int i = find(3, l, noWarning);
    return i < 0 ? nullptr : getVectorData(i);
  }
  void clear() { detail::ok(lsg_domain_point_data_clear(h_), "PointData::clear"); }
  bool empty() const { return getScalarDataSize() == 0 && getVectorDataSize() == 0; }

private:
  void put(const T *p, size_t n, int comps, const std::string &label, int replace) {
    detail::ok(lsg_domain_point_data_insert(h_, label.c_str(), comps, p, (int64_t)n, replace), "PointData::insert");
  }
  std::string label(int comps, int i) const {
    char buf[1024] = {0};
    const int n = lsg_domain_point_data_label(h_, comps, i, buf, sizeof(buf));
    if (n < 0) detail::ok(n, "PointData::label");
    return buf;
  }
  int find(int comps, const std::string &l, bool noWarning) const {
    const int n = lsg_domain_point_data_count(h_, comps);
    for (int i = 0; i < n; ++i)
      if (label(comps, i) == l) return i;
    if (!noWarning) std::fprintf(stderr, "[viennals_gpu] PointData: no data with label %s\n", l.c_str());
    return -1;
  }
  bool fetch(int comps, int i, T *, int, std::vector<T> *out) const {
    const int64_t n = lsg_domain_point_data_size(h_, comps, i);
    if (n < 0) return false;
    out->resize((size_t)n * comps);
    return detail::ok(lsg_domain_point_data_get(h_, comps, i, out->data(), n), "PointData::get");
  }
};

template <class T, int D> class Domain {
  static_assert(std::is_same<T, double>::value, "the GPU path computes in double precision");
  static_assert(D == 2 || D == 3, "D must be 2 or 3");
  lsg_domain *h_ = nullptr;
  double gridDelta_ = 1.0;
  PointData<T> pointData_{nullptr};

public:
  using BoundaryType = BoundaryConditionEnum;
  using PointDataType = PointData<T>;
  using PointValueVectorType = std::vector<std::pair<std::array<int32_t, D>, T>>;
  static constexpr T POS_VALUE = 1.7976931348623157e308;
  static constexpr T NEG_VALUE = -1.7976931348623157e308;

  // Domain(gridDelta): all boundaries infinite (lsDomain.hpp:55-66)
  explicit Domain(double gridDelta = 1.0) : gridDelta_(gridDelta) {
    int32_t lo[D] = {}, hi[D] = {};
    int bc[D];
    for (int i = 0; i < D; ++i) bc[i] = LSG_INFINITE;
    detail::ok(lsg_domain_create(D, lo, hi, bc, gridDelta, &h_), "Domain");
    pointData_ = PointData<T>(h_);
  }
  // Domain(bounds, boundaryConditions, gridDelta) (lsDomain.hpp:68-78)
  Domain(const double *bounds, const BoundaryType *bcs, double gridDelta = 1.0) : gridDelta_(gridDelta) {
    int32_t lo[D], hi[D];
    int bc[D];
    for (int i = 0; i < D; ++i) {
      lo[i] = (int32_t)std::floor(bounds[2 * i] / gridDelta);
      hi[i] = (int32_t)std::ceil(bounds[2 * i + 1] / gridDelta);
      bc[i] = (int)bcs[i];
    }
    detail::ok(lsg_domain_create(D, lo, hi, bc, gridDelta, &h_), "Domain");
    pointData_ = PointData<T>(h_);
  }
  ~Domain() {
    if (h_) lsg_domain_destroy(h_);
  }
  Domain(const Domain &) = delete;
  Domain &operator=(const Domain &) = delete;

  template <class... Args> static auto New(Args &&...args) {
    return SmartPointer<Domain>::New(std::forward<Args>(args)...);
  }

  lsg_domain *handle() const { return h_; }
  double getGridDelta() const { return gridDelta_; }
  // Domain::getPointData (lsDomain.hpp:166-168)
  PointData<T> &getPointData() { return pointData_; }

  // Domain::deepCopy (lsDomain.hpp:124-129)
  void deepCopy(const SmartPointer<Domain> &other) { detail::ok(lsg_domain_deep_copy(h_, other->h_), "deepCopy"); }
  // Domain::insertPoints (lsDomain.hpp:135-138)
  void insertPoints(const PointValueVectorType &pts, bool = true) {
    std::vector<int32_t> idx;
    std::vector<double> val;
    idx.reserve(pts.size() * D);
    for (auto &p : pts) {
      for (int i = 0; i < D; ++i) idx.push_back(p.first[i]);
      val.push_back(p.second);
    }
    detail::ok(lsg_domain_insert_points(h_, (int64_t)val.size(), idx.data(), val.data(), 2), "insertPoints");
  }
  unsigned getNumberOfPoints() const {
    int64_t n = 0;
    detail::ok(lsg_domain_num_points(h_, &n), "getNumberOfPoints");
    return (unsigned)n;
  }
  int getLevelSetWidth() const {
    int w = 0;
    detail::ok(lsg_domain_width(h_, &w), "getLevelSetWidth");
    return w;
  }
  // Run-start list (HRLE iteration order); undefined runs carry +-POS_VALUE.
  void getRuns(std::vector<std::array<int32_t, D>> &idx, std::vector<T> &val) const {
    int64_t n = 0;
    if (!detail::ok(lsg_domain_num_runs(h_, &n), "getRuns")) return;
    std::vector<int32_t> flat((size_t)n * D);
    val.resize((size_t)n);
    detail::ok(lsg_domain_get_runs(h_, flat.data(), val.data(), n), "getRuns");
    idx.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < D; ++k) idx[i][k] = flat[i * D + k];
  }
};

// ---------------------------------------------------------- geometries ---

template <class T, int D> struct Sphere {
  std::array<T, 3> origin{};
  T radius;
  Sphere(const T *o, T r) : radius(r) {
    for (int i = 0; i < D; ++i) origin[i] = o[i];
  }
  template <class... A> static auto New(A &&...a) { return SmartPointer<Sphere>::New(std::forward<A>(a)...); }
};
template <class T, int D> struct Plane {
  std::array<T, 3> origin{}, normal{};
  Plane(const T *o, const T *n) {
    for (int i = 0; i < D; ++i) origin[i] = o[i], normal[i] = n[i];
  }
  template <class... A> static auto New(A &&...a) { return SmartPointer<Plane>::New(std::forward<A>(a)...); }
};
template <class T, int D> struct Box {
  std::array<T, 3> minCorner{}, maxCorner{};
  Box(const T *mn, const T *mx) {
    for (int i = 0; i < D; ++i) minCorner[i] = mn[i], maxCorner[i] = mx[i];
  }
  template <class... A> static auto New(A &&...a) { return SmartPointer<Box>::New(std::forward<A>(a)...); }
};

// MarkVoidPoints with its default options (lsMarkVoidPoints.hpp:156-331).
template <class T, int D> class MarkVoidPoints {
  SmartPointer<Domain<T, D>> ls_;

public:
  static constexpr char voidPointLabel[] = "VoidPointMarkers";
  MarkVoidPoints() = default;
  explicit MarkVoidPoints(SmartPointer<Domain<T, D>> ls) : ls_(ls) {}
  void setLevelSet(SmartPointer<Domain<T, D>> ls) { ls_ = ls; }
  void apply() { detail::ok(lsg_mark_void_points(ls_->handle()), "MarkVoidPoints"); }
};

// CompareChamfer (lsCompareChamfer.hpp:104-330) on the GPU surface points.
template <class T, int D> class CompareChamfer {
  SmartPointer<Domain<T, D>> target_, sample_;
  double r_[5] = {0, 0, 0, 0, 0};
  int64_t n_[2] = {0, 0};

public:
  CompareChamfer() = default;
  CompareChamfer(SmartPointer<Domain<T, D>> t, SmartPointer<Domain<T, D>> s) : target_(t), sample_(s) {}
  void setLevelSetTarget(SmartPointer<Domain<T, D>> t) { target_ = t; }
  void setLevelSetSample(SmartPointer<Domain<T, D>> s) { sample_ = s; }
  void apply() { detail::ok(lsg_compare_chamfer(target_->handle(), sample_->handle(), r_, n_), "CompareChamfer"); }
  T getForwardDistance() const { return r_[0]; }
  T getBackwardDistance() const { return r_[1]; }
  T getChamferDistance() const { return r_[2]; }
  T getRMSChamferDistance() const { return r_[3]; }
  T getMaxDistance() const { return r_[4]; }
  unsigned getNumTargetPoints() const { return (unsigned)n_[0]; }
  unsigned getNumSamplePoints() const { return (unsigned)n_[1]; }
};

// CompareVolume (lsCompareVolume.hpp:176-289) on the GPU (CompareArea in 2D).
template <class T, int D> class CompareVolume {
  SmartPointer<Domain<T, D>> target_, sample_;
  uint64_t cells_ = 0;
  double volume_ = 0;

public:
  CompareVolume() = default;
  CompareVolume(SmartPointer<Domain<T, D>> t, SmartPointer<Domain<T, D>> s) : target_(t), sample_(s) {}
  void setLevelSetTarget(SmartPointer<Domain<T, D>> t) { target_ = t; }
  void setLevelSetSample(SmartPointer<Domain<T, D>> s) { sample_ = s; }
  void apply() { detail::ok(lsg_compare_volume(target_->handle(), sample_->handle(), &cells_, &volume_), "CompareVolume"); }
  double getVolumeMismatch() const { return volume_; }
  double getAreaMismatch() const { return volume_; }
  unsigned long getCellCount() const { return (unsigned long)cells_; }
};

template <class T, int D> class MakeGeometry {
  SmartPointer<Domain<T, D>> ls_;
  int kind_ = -1;
  std::array<T, 3> a_{}, b_{};
  T r_ = 0;
  int ignoreMask_ = 0;

public:
  MakeGeometry(SmartPointer<Domain<T, D>> ls, SmartPointer<Sphere<T, D>> s) : ls_(ls), kind_(0), a_(s->origin), r_(s->radius) {}
  MakeGeometry(SmartPointer<Domain<T, D>> ls, SmartPointer<Plane<T, D>> p) : ls_(ls), kind_(1), a_(p->origin), b_(p->normal) {}
  MakeGeometry(SmartPointer<Domain<T, D>> ls, SmartPointer<Box<T, D>> b) : ls_(ls), kind_(2), a_(b->minCorner), b_(b->maxCorner) {}
  // lsMakeGeometry.hpp:125-141: spheres keep periodic indices unmapped, boxes
  // drop their boundary elements along periodic axes; planes ignore it
  void setIgnoreBoundaryConditions(bool ignore) { ignoreMask_ = ignore ? 7 : 0; }
  template <std::size_t N> void setIgnoreBoundaryConditions(std::array<bool, N> ignore) {
    ignoreMask_ = 0;
    for (unsigned i = 0; i < D && i < N; ++i) ignoreMask_ |= ignore[i] ? (1 << i) : 0;
  }
  void apply() {
    if (kind_ == 0) detail::ok(lsg_make_sphere_ignore_bc(ls_->handle(), a_.data(), r_, ignoreMask_), "MakeGeometry");
    else if (kind_ == 1) detail::ok(lsg_make_plane(ls_->handle(), a_.data(), b_.data()), "MakeGeometry");
    else if (kind_ == 2) detail::ok(lsg_make_box_ignore_bc(ls_->handle(), a_.data(), b_.data(), ignoreMask_), "MakeGeometry");
  }
};

// ------------------------------------------------------ band maintenance ---

template <class T, int D> class Expand {
  SmartPointer<Domain<T, D>> ls_;
  int width_ = 0;

public:
  Expand() = default;
  Expand(SmartPointer<Domain<T, D>> ls, int width = 0) : ls_(ls), width_(width) {}
  void setLevelSet(SmartPointer<Domain<T, D>> ls) { ls_ = ls; }
  void setWidth(int w) { width_ = w; }
  void apply() {
    if (!ls_) {
      std::fprintf(stderr, "[viennals_gpu] No level set passed to Expand. Not expanding.\n");
      return;
    }
    detail::ok(lsg_expand(ls_->handle(), width_), "Expand");
  }
};

template <class T, int D> class Reduce {
  SmartPointer<Domain<T, D>> ls_;
  int width_ = 0;

public:
  Reduce() = default;
  Reduce(SmartPointer<Domain<T, D>> ls, int width = 0, bool = false) : ls_(ls), width_(width) {}
  void setLevelSet(SmartPointer<Domain<T, D>> ls) { ls_ = ls; }
  void setWidth(int w) { width_ = w; }
  void apply() {
    if (!ls_) {
      std::fprintf(stderr, "[viennals_gpu] No level set was passed to Reduce.\n");
      return;
    }
    detail::ok(lsg_reduce(ls_->handle(), width_), "Reduce");
  }
};

template <class T, int D> class Prune {
  SmartPointer<Domain<T, D>> ls_;
  bool rsz_ = false;

public:
  Prune() = default;
  explicit Prune(SmartPointer<Domain<T, D>> ls) : ls_(ls) {}
  void setLevelSet(SmartPointer<Domain<T, D>> ls) { ls_ = ls; }
  void setRemoveStrayZeros(bool v) { rsz_ = v; }
  void apply() {
    if (!ls_) {
      std::fprintf(stderr, "[viennals_gpu] No level set was passed to Prune.\n");
      return;
    }
    detail::ok(lsg_prune(ls_->handle(), rsz_ ? 1 : 0), "Prune");
  }
};

template <class T, int D> class BooleanOperation {
  SmartPointer<Domain<T, D>> a_, b_;
  BooleanOperationEnum op_ = BooleanOperationEnum::INTERSECT;

public:
  BooleanOperation() = default;
  BooleanOperation(SmartPointer<Domain<T, D>> a, SmartPointer<Domain<T, D>> b,
                   BooleanOperationEnum op = BooleanOperationEnum::INTERSECT)
      : a_(a), b_(b), op_(op) {}
  void setLevelSet(SmartPointer<Domain<T, D>> a) { a_ = a; }
  void setSecondLevelSet(SmartPointer<Domain<T, D>> b) { b_ = b; }
  void setBooleanOperation(BooleanOperationEnum op) { op_ = op; }
  void apply() {
    if (!a_ || !b_) {
      std::fprintf(stderr, "[viennals_gpu] BooleanOperation needs two level sets.\n");
      return;
    }
    detail::ok(lsg_boolean(a_->handle(), b_->handle(), (int)op_), "BooleanOperation");
  }
};

// ------------------------------------------------------ velocity fields ---

// Abstract velocity field (lsVelocityField.hpp:11-41).  The GPU path evaluates
// velocities on the device; only BuiltinVelocityField is accepted by Advect.
template <class T> class VelocityField {
public:
  virtual T getScalarVelocity(const Vec3D<T> &, int, const Vec3D<T> &, unsigned long) { return 0; }
  virtual Vec3D<T> getVectorVelocity(const Vec3D<T> &, int, const Vec3D<T> &, unsigned long) { return {0, 0, 0}; }
  virtual T getDissipationAlpha(int, int, const Vec3D<T> &) { return 0; }
  virtual ~VelocityField() = default;
};

// scalar = b                                          (kind LSG_VEL_CONSTANT)
//        = b + (n[dir] > 0 ? aPos : aNeg) * |n[dir]|^k (kind LSG_VEL_NORMAL_POWER)
//        = b + aPos * (|n[0]| + |n[1]| + |n[2]|)     (kind LSG_VEL_NORMAL_SUM)
// vector = vec
template <class T> class BuiltinVelocityField : public VelocityField<T> {
public:
  int kind = LSG_VEL_CONSTANT, dir = 2, k = 1;
  T b = 0, aPos = 0, aNeg = 0;
  Vec3D<T> vec{0, 0, 0};
  Vec3D<T> alpha{0, 0, 0};  // getDissipationAlpha: constant per direction (LLFA)
  static SmartPointer<BuiltinVelocityField> constant(T scalar, Vec3D<T> vector = {0, 0, 0}) {
    auto v = SmartPointer<BuiltinVelocityField>::New();
    v->b = scalar;
    v->vec = vector;
    return v;
  }
  static SmartPointer<BuiltinVelocityField> normalPower(int dir, int k, T b, T aPos, T aNeg = 0,
                                                       Vec3D<T> vector = {0, 0, 0}) {
    auto v = SmartPointer<BuiltinVelocityField>::New();
    v->kind = LSG_VEL_NORMAL_POWER;
    v->dir = dir;
    v->k = k;
    v->b = b;
    v->aPos = aPos;
    v->aNeg = aNeg;
    v->vec = vector;
    return v;
  }
  static SmartPointer<BuiltinVelocityField> normalSum(T a, T b = 0, Vec3D<T> vector = {0, 0, 0}) {
    auto v = SmartPointer<BuiltinVelocityField>::New();
    v->kind = LSG_VEL_NORMAL_SUM;
    v->b = b;
    v->aPos = a;
    v->vec = vector;
    return v;
  }
  T getScalarVelocity(const Vec3D<T> &, int, const Vec3D<T> &n, unsigned long) override {
    if (kind == LSG_VEL_CONSTANT) return b;
    if (kind == LSG_VEL_NORMAL_SUM) return b + aPos * (std::fabs(n[0]) + std::fabs(n[1]) + std::fabs(n[2]));
    T c = n[dir], m = 1;
    if (k >= 1) {
      m = std::fabs(c);
      for (int j = 1; j < k; ++j) m = m * std::fabs(c);
    }
    return b + ((c > 0) ? aPos : aNeg) * m;
  }
  Vec3D<T> getVectorVelocity(const Vec3D<T> &, int, const Vec3D<T> &, unsigned long) override { return vec; }
  T getDissipationAlpha(int direction, int, const Vec3D<T> &) override { return alpha[direction]; }
};

// Built-in functors per material (several level sets: the material of a
// point is the index of the level set below or at it, lsAdvect.hpp:513-527).
// Material m uses perMaterial[m]; later materials use the last entry.
template <class T> class MaterialVelocityField : public VelocityField<T> {
public:
  std::vector<SmartPointer<BuiltinVelocityField<T>>> perMaterial;
  explicit MaterialVelocityField(std::vector<SmartPointer<BuiltinVelocityField<T>>> v = {})
      : perMaterial(std::move(v)) {}
  BuiltinVelocityField<T> *of(int m) const {
    if (perMaterial.empty()) return nullptr;
    return perMaterial[(m >= 0 && m < (int)perMaterial.size()) ? m : (int)perMaterial.size() - 1].get();
  }
  T getScalarVelocity(const Vec3D<T> &c, int m, const Vec3D<T> &n, unsigned long id) override {
    auto *f = of(m);
    return f ? f->getScalarVelocity(c, m, n, id) : T(0);
  }
  Vec3D<T> getVectorVelocity(const Vec3D<T> &c, int m, const Vec3D<T> &n, unsigned long id) override {
    auto *f = of(m);
    return f ? f->getVectorVelocity(c, m, n, id) : Vec3D<T>{0, 0, 0};
  }
};

// --------------------------------------------------------------- Advect ---

template <class T, int D> class Advect {
  lsg_advect *h_ = nullptr;
  std::vector<SmartPointer<Domain<T, D>>> levelSets_;
  SmartPointer<VelocityField<T>> velocities_;
  std::function<bool(SmartPointer<Domain<T, D>>)> velocityUpdateCallback_;
  std::exception_ptr callbackError_;

  // Between RK stages: the user's callback with the advected level set, then
  // the velocity field again, so built-in functors it changed take effect.
  // An exception is held here (it must not cross the C library), the
  // library stops the apply at that stage, and apply() rethrows it.
  static int velocityUpdate(lsg_domain *, void *user) {
    auto *self = static_cast<Advect *>(user);
    try {
      (void)self->velocityUpdateCallback_(self->levelSets_.back());
      (void)self->pushVelocity();
    } catch (...) {
      if (!self->callbackError_) self->callbackError_ = std::current_exception();
      return -1;  // the library stops the apply at this stage
    }
    return 0;
  }

  // Any other VelocityField subclass is evaluated on the host through its
  // virtual methods, called back for every active point of every stage.
  static double hostScalar(const double *c, int material, const double *n, uint64_t id, void *user) {
    auto *v = static_cast<VelocityField<T> *>(user);
    return (double)v->getScalarVelocity(Vec3D<T>{(T)c[0], (T)c[1], (T)c[2]}, material,
                                        Vec3D<T>{(T)n[0], (T)n[1], (T)n[2]}, (unsigned long)id);
  }
  static void hostVector(const double *c, int material, const double *n, uint64_t id, void *user, double *out) {
    auto *v = static_cast<VelocityField<T> *>(user);
    const Vec3D<T> r = v->getVectorVelocity(Vec3D<T>{(T)c[0], (T)c[1], (T)c[2]}, material,
                                            Vec3D<T>{(T)n[0], (T)n[1], (T)n[2]}, (unsigned long)id);
    out[0] = (double)r[0];
    out[1] = (double)r[1];
    out[2] = (double)r[2];
  }

  static double hostAlpha(int direction, int material, const double *cd, void *user) {
    auto *v = static_cast<VelocityField<T> *>(user);
    return (double)v->getDissipationAlpha(direction, material, Vec3D<T>{(T)cd[0], (T)cd[1], (T)cd[2]});
  }

  bool pushVelocity() {
    if (!velocities_) {
      std::fprintf(stderr, "[viennals_gpu] No velocity field passed to Advect. Not advecting.\n");
      return false;
    }
    if (auto *mv = dynamic_cast<MaterialVelocityField<T> *>(velocities_.get())) {
      if (mv->perMaterial.empty() || mv->perMaterial.size() > 8) {
        std::fprintf(stderr, "[viennals_gpu] MaterialVelocityField needs 1 to 8 functors.\n");
        return false;
      }
      auto *d = mv->perMaterial.back().get();
      double vd[3] = {d->vec[0], d->vec[1], d->vec[2]};
      double ad[3] = {d->alpha[0], d->alpha[1], d->alpha[2]};
      if (!detail::ok(lsg_advect_set_velocity_builtin(h_, d->kind, d->dir, d->k, d->b, d->aPos, d->aNeg, vd),
                      "setVelocityField") ||
          !detail::ok(lsg_advect_set_dissipation_alpha_builtin(h_, -1, ad), "setVelocityField"))
        return false;
      for (size_t m = 0; m < mv->perMaterial.size(); ++m) {
        auto *f = mv->perMaterial[m].get();
        double vf[3] = {f->vec[0], f->vec[1], f->vec[2]};
        if (!detail::ok(lsg_advect_set_velocity_builtin_material(h_, (int)m, f->kind, f->dir, f->k, f->b, f->aPos,
                                                                 f->aNeg, vf),
                        "setVelocityField"))
          return false;
        double af[3] = {f->alpha[0], f->alpha[1], f->alpha[2]};
        if (!detail::ok(lsg_advect_set_dissipation_alpha_builtin(h_, (int)m, af), "setVelocityField")) return false;
      }
      return true;
    }
    auto *b = dynamic_cast<BuiltinVelocityField<T> *>(velocities_.get());
    if (!b)
      return detail::ok(lsg_advect_set_velocity_host(h_, &Advect::hostScalar, &Advect::hostVector, velocities_.get()),
                        "setVelocityField") &&
             detail::ok(lsg_advect_set_dissipation_alpha_host(h_, &Advect::hostAlpha, velocities_.get()),
                        "setVelocityField");
    double vec[3] = {b->vec[0], b->vec[1], b->vec[2]};
    double al[3] = {b->alpha[0], b->alpha[1], b->alpha[2]};
    return detail::ok(lsg_advect_set_velocity_builtin(h_, b->kind, b->dir, b->k, b->b, b->aPos, b->aNeg, vec),
                      "setVelocityField") &&
           detail::ok(lsg_advect_set_dissipation_alpha_builtin(h_, -1, al), "setVelocityField");
  }

public:
  static constexpr char velocityLabel[] = "AdvectionVelocities";
  static constexpr char dissipationLabel[] = "Dissipation";

  Advect() { detail::ok(lsg_advect_create(&h_), "Advect"); }
  explicit Advect(SmartPointer<Domain<T, D>> ls) : Advect() { insertNextLevelSet(ls); }
  Advect(SmartPointer<Domain<T, D>> ls, SmartPointer<VelocityField<T>> v) : Advect() {
    insertNextLevelSet(ls);
    setVelocityField(v);
  }
  Advect(std::vector<SmartPointer<Domain<T, D>>> lss, SmartPointer<VelocityField<T>> v) : Advect() {
    for (auto &l : lss) insertNextLevelSet(l);
    setVelocityField(v);
  }
  ~Advect() {
    if (h_) lsg_advect_destroy(h_);
  }
  Advect(const Advect &) = delete;
  Advect &operator=(const Advect &) = delete;

  void insertNextLevelSet(SmartPointer<Domain<T, D>> ls) {
    levelSets_.push_back(ls);
    detail::ok(lsg_advect_insert_next_level_set(h_, ls->handle()), "insertNextLevelSet");
  }
  void clearLevelSets() {
    levelSets_.clear();
    detail::ok(lsg_advect_clear_level_sets(h_), "clearLevelSets");
  }
  void setVelocityField(SmartPointer<VelocityField<T>> v) { velocities_ = v; }
  void setAdvectionTime(double t) { detail::ok(lsg_advect_set_advection_time(h_, t), "setAdvectionTime"); }
  void setSingleStep(bool s) { detail::ok(lsg_advect_set_single_step(h_, s ? 1 : 0), "setSingleStep"); }
  void setTimeStepRatio(const double &cfl) { detail::ok(lsg_advect_set_time_step_ratio(h_, cfl), "setTimeStepRatio"); }
  void setCalculateNormalVectors(bool c) {
    detail::ok(lsg_advect_set_calculate_normal_vectors(h_, c ? 1 : 0), "setCalculateNormalVectors");
  }
  void setIgnoreVoids(bool v) { detail::ok(lsg_advect_set_ignore_voids(h_, v ? 1 : 0), "setIgnoreVoids"); }
  void setAdaptiveTimeStepping(bool on = true, unsigned sub = 20) {
    detail::ok(lsg_advect_set_adaptive_time_stepping(h_, on ? 1 : 0, sub), "setAdaptiveTimeStepping");
  }
  // Advect::setVelocityUpdateCallback (lsAdvect.hpp:966-969)
  void setVelocityUpdateCallback(std::function<bool(SmartPointer<Domain<T, D>>)> callback) {
    velocityUpdateCallback_ = std::move(callback);
    detail::ok(lsg_advect_set_velocity_update_callback(h_, velocityUpdateCallback_ ? &Advect::velocityUpdate : nullptr,
                                                       this),
               "setVelocityUpdateCallback");
  }
  void setSaveAdvectionVelocities(bool s) {
    detail::ok(lsg_advect_set_save_advection_velocities(h_, s ? 1 : 0), "setSaveAdvectionVelocities");
  }
  void setSpatialScheme(SpatialSchemeEnum s) { detail::ok(lsg_advect_set_spatial_scheme(h_, (int)s), "setSpatialScheme"); }
  [[deprecated("Use setSpatialScheme instead")]] void setIntegrationScheme(IntegrationSchemeEnum s) { setSpatialScheme(s); }
  void setTemporalScheme(TemporalSchemeEnum s) {
    detail::ok(lsg_advect_set_temporal_scheme(h_, (int)s), "setTemporalScheme");
  }
  void setDissipationAlpha(const double &a) { detail::ok(lsg_advect_set_dissipation_alpha(h_, a), "setDissipationAlpha"); }
  void setCheckDissipation(bool c) { detail::ok(lsg_advect_set_check_dissipation(h_, c ? 1 : 0), "setCheckDissipation"); }
  void setUpdatePointData(bool u) { detail::ok(lsg_advect_set_update_point_data(h_, u ? 1 : 0), "setUpdatePointData"); }

  double getAdvectedTime() const {
    double t = 0;
    lsg_advect_get_advected_time(h_, &t);
    return t;
  }
  double getCurrentTimeStep() const {
    double t = 0;
    lsg_advect_get_current_time_step(h_, &t);
    return t;
  }
  unsigned getNumberOfTimeSteps() const {
    unsigned n = 0;
    lsg_advect_get_number_of_time_steps(h_, &n);
    return n;
  }

  void prepareLS() {
    if (levelSets_.empty()) {
      std::fprintf(stderr, "[viennals_gpu] No level sets passed to Advect.\n");
      return;
    }
    detail::ok(lsg_advect_prepare_ls(h_), "prepareLS");
  }

  // Advect::apply (lsAdvect.hpp:1017-1043)
  void apply() {
    if (levelSets_.empty()) {
      std::fprintf(stderr, "[viennals_gpu] No level sets passed to Advect. Not advecting.\n");
      return;
    }
    if (velocities_ == nullptr) {
      std::fprintf(stderr, "[viennals_gpu] No velocity field passed to Advect. Not advecting.\n");
      return;
    }
    if (!pushVelocity()) return;
    callbackError_ = nullptr;
    const int status = lsg_advect_apply(h_);
    if (callbackError_) std::rethrow_exception(std::exchange(callbackError_, nullptr));
    detail::ok(status, "Advect::apply");
  }
};

}  // namespace viennals
