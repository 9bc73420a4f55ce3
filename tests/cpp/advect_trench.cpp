// Uses the viennals::Advect<T,D> facade exactly as reference code would
// (compare examples/SquareEtch and tests/Advection): builds a trench with
// MakeGeometry + BooleanOperation, advects it with WENO5 + RK3, and writes
// the run lists before and after to binary files for the parity test.
// With "host" as third argument the velocity is a user-defined
// VelocityField subclass (evaluated on the host through its virtual
// methods) computing the same function as the built-in functor.  With
// "callback" a velocity-update callback lowers the built-in functor's
// constant by 0.01 at every call (between the RK stages).
//   usage: advect_trench <out_prefix> <time> [builtin|host|callback]
#include <viennals_gpu/lsAdvect.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

namespace ls = viennals;
using T = double;
constexpr int D = 3;

static void dump(const ls::SmartPointer<ls::Domain<T, D>> &d, const std::string &path) {
  std::vector<std::array<int32_t, D>> idx;
  std::vector<T> val;
  d->getRuns(idx, val);
  FILE *f = std::fopen(path.c_str(), "wb");
  int64_t n = (int64_t)val.size();
  std::fwrite(&n, sizeof(n), 1, f);
  std::fwrite(idx.data(), sizeof(int32_t) * D, n, f);
  std::fwrite(val.data(), sizeof(T), n, f);
  std::fclose(f);
}

// The directional etch velocity -(max(0, n_z) + 0.1) written as reference
// code would write it (same operations as the built-in normal-power functor).
class EtchVelocity : public ls::VelocityField<T> {
public:
  T getScalarVelocity(const ls::Vec3D<T> &, int, const ls::Vec3D<T> &n, unsigned long) override {
    const T c = n[2];
    const T m = std::fabs(c);
    return -0.1 + ((c > 0) ? -1.0 : 0.0) * m;
  }
};

int main(int argc, char **argv) {
  std::string prefix = argc > 1 ? argv[1] : "trench";
  double time = argc > 2 ? std::atof(argv[2]) : 2.0;
  double bounds[2 * D] = {-32, 32, -32, 32, -32, 32};
  ls::Domain<T, D>::BoundaryType bcs[D] = {ls::BoundaryConditionEnum::REFLECTIVE_BOUNDARY,
                                           ls::BoundaryConditionEnum::REFLECTIVE_BOUNDARY,
                                           ls::BoundaryConditionEnum::INFINITE_BOUNDARY};
  auto substrate = ls::Domain<T, D>::New(bounds, bcs, 1.0);
  T origin[3] = {0, 0, 0}, normal[3] = {0, 0, 1};
  ls::MakeGeometry<T, D>(substrate, ls::Plane<T, D>::New(origin, normal)).apply();
  auto trench = ls::Domain<T, D>::New(bounds, bcs, 1.0);
  T mn[3] = {-8, -8, -16}, mx[3] = {8, 8, 1};
  ls::MakeGeometry<T, D>(trench, ls::Box<T, D>::New(mn, mx)).apply();
  ls::BooleanOperation<T, D>(substrate, trench, ls::BooleanOperationEnum::RELATIVE_COMPLEMENT).apply();
  dump(substrate, prefix + "_initial.bin");

  const bool host = argc > 3 && std::string(argv[3]) == "host";
  ls::SmartPointer<ls::VelocityField<T>> velocity;
  if (host) velocity = ls::SmartPointer<EtchVelocity>::New();
  else velocity = ls::BuiltinVelocityField<T>::normalPower(2, 1, -0.1, -1.0, 0.0);
  ls::Advect<T, D> advect;
  advect.insertNextLevelSet(substrate);
  advect.setVelocityField(velocity);
  advect.setSpatialScheme(ls::SpatialSchemeEnum::WENO_5TH_ORDER);
  advect.setTemporalScheme(ls::TemporalSchemeEnum::RUNGE_KUTTA_3RD_ORDER);
  advect.setTimeStepRatio(0.4);
  advect.setAdvectionTime(time);
  unsigned calls = 0;
  if (argc > 3 && std::string(argv[3]) == "callback") {
    auto builtin = std::dynamic_pointer_cast<ls::BuiltinVelocityField<T>>(velocity);
    advect.setVelocityUpdateCallback([builtin, &calls](ls::SmartPointer<ls::Domain<T, D>> top) {
      ++calls;
      builtin->b -= 0.01;
      return top != nullptr;
    });
  }
  advect.apply();
  if (calls) std::printf("callbacks %u\n", calls);
  dump(substrate, prefix + "_final.bin");
  std::printf("steps %u advected %.17g points %u width %d\n", advect.getNumberOfTimeSteps(),
              advect.getAdvectedTime(), substrate->getNumberOfPoints(), substrate->getLevelSetWidth());
  return 0;
}
