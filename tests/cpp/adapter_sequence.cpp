// The forming driver's calls on the device nonlinear system, through the C++
// adapter of include/sparlab_gpu_system.hpp, in the order of the reference's
// FormingAnalysis::run (src/fem/Forming.cpp): a restart from a saved history
// (set_committed / set_internal, :693-694), plastic() and averaged(e)
// (:712-714), per increment the tangent evaluation (:770), symmetric() for the
// factorisation (:1005), a residual-only evaluation of the line search
// (:779), commit (:1202) and max_plastic_strain (:1211); at the end committed(e),
// averaged(e), internal(e), elastoplastic(e) and the stress recovery
// (:1263-1281), and committed_all / internal_all (:1339-1340), which restart a
// second system.
//
// Test driver (tests/test_adapter_gpu.py runs it): reads the mesh and the
// displacement of each increment from argv[1], writes what the calls
// returned to argv[2] (raw little-endian doubles, in the order below), exits 0
// when its own consistency checks pass.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "sparlab_gpu_system.hpp"

namespace {

template <class T>
void read(std::ifstream& in, T* p, std::size_t n) {
  in.read(reinterpret_cast<char*>(p), static_cast<std::streamsize>(n * sizeof(T)));
  if (!in) throw std::runtime_error("short input file");
}

struct Out {
  std::vector<double> v;
  void put(double x) { v.push_back(x); }
  void put(const std::vector<double>& x) { v.insert(v.end(), x.begin(), x.end()); }
};

// The reference's 39-double PlasticState record: plastic strain 6, beta 6,
// alpha_1..4 24, equivalent plastic strain, thickness strain, loading.
void put_state(Out& o, const sparlab_gpu::PlasticState& p) {
  for (double x : p.plastic_strain) o.put(x);
  for (double x : p.back_stress) o.put(x);
  for (const auto& a : p.back_stresses)
    for (double x : a) o.put(x);
  o.put(p.equivalent_plastic_strain);
  o.put(p.thickness_strain);
  o.put(p.loading ? 1.0 : 0.0);
}

sparlab_material material(int which) {
  sparlab_material m{};
  m.youngs_modulus = 70e9;
  m.poisson_ratio = 0.33;
  m.yield_stress = 95e6;
  m.saturation_stress = 150e6;
  m.saturation_rate = 10.0;
  m.hill_shear_lm[0] = m.hill_shear_lm[1] = 1.5;
  m.rolling_direction[0] = 1.0;
  m.sheet_normal[2] = 1.0;
  if (which == 1) {  // Hill48 r 0.7 / 0.6 / 0.8 and two Chaboche backstresses: a non-symmetric tangent
    m.criterion = 1;
    m.hill_r[0] = 0.7;
    m.hill_r[1] = 0.6;
    m.hill_r[2] = 0.8;
    m.num_backstresses = 2;
    m.backstress_modulus[0] = 6000e6;
    m.backstress_recovery[0] = 60.0;
    m.backstress_modulus[1] = 1500e6;
    m.backstress_recovery[1] = 10.0;
  }
  return m;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: adapter_sequence input output\n");
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    int64_t head[5];
    read(in, head, 5);
    const int64_t nn = head[0], ne = head[1], steps = head[2], which = head[3], thickness = head[4];
    std::vector<double> coords(static_cast<std::size_t>(3 * nn));
    std::vector<int32_t> conn(static_cast<std::size_t>(8 * ne));
    read(in, coords.data(), coords.size());
    read(in, conn.data(), conn.size());
    std::vector<std::vector<double>> u(static_cast<std::size_t>(steps), std::vector<double>(coords.size()));
    for (auto& uk : u) read(in, uk.data(), uk.size());

    const sparlab_rule rule{2, static_cast<int32_t>(thickness), 2};
    sparlab_gpu::System system(coords, conn, material(static_cast<int>(which)), rule, SPARLAB_FINITE_LOGARITHMIC);
    const int points = system.info().points;
    int failures = 0;
    auto expect = [&failures](bool ok, const char* what) {
      if (!ok) {
        std::fprintf(stderr, "check failed: %s\n", what);
        ++failures;
      }
    };
    // a restart from a saved (here virgin) history, as a FormingStart carries it
    system.set_committed(std::vector<std::vector<sparlab_gpu::PlasticState>>(
        static_cast<std::size_t>(ne), std::vector<sparlab_gpu::PlasticState>(static_cast<std::size_t>(points))));
    system.set_internal(std::vector<std::array<double, 9>>(static_cast<std::size_t>(ne)));
    Out out;
    out.put(system.plastic() ? 1.0 : 0.0);
    out.put(system.symmetric() ? 1.0 : 0.0);
    bool any_averaged = false, all_plastic = true;
    for (int64_t e = 0; e < ne; ++e) {
      any_averaged = any_averaged || system.averaged(e);
      all_plastic = all_plastic && system.elastoplastic(e);
    }
    out.put(any_averaged ? 1.0 : 0.0);
    out.put(all_plastic ? 1.0 : 0.0);
    for (int64_t k = 0; k < steps; ++k) {
      const sparlab_gpu::Evaluation ev = system.evaluate(u[static_cast<std::size_t>(k)], 0.0, true);
      const bool symmetric_factor = system.symmetric();  // and the contact's flag, none here
      const sparlab_gpu::Evaluation line = system.evaluate(u[static_cast<std::size_t>(k)], 0.0, false);
      expect(line.internal == ev.internal, "the residual-only evaluation repeats f_int bit for bit");
      expect(line.energy == ev.energy && line.yielding_points == ev.yielding_points,
             "the residual-only evaluation repeats energy and yielding");
      double abs_sum = 0.0;
      for (double x : ev.tangent_values) abs_sum += std::fabs(x);
      expect(ev.tangent_values.size() == system.inner().size(), "one tangent value per pattern entry");
      system.commit();
      const double top = system.max_plastic_strain();
      out.put(ev.energy);
      out.put(ev.gross);
      out.put(static_cast<double>(ev.yielding_points));
      out.put(top);
      out.put(symmetric_factor ? 1.0 : 0.0);
      out.put(abs_sum);
      out.put(ev.internal);
    }
    const auto all = system.committed_all();
    const auto modes = system.internal_all();
    double top = 0.0;
    for (int64_t e = 0; e < ne; ++e) {
      const auto one = system.committed(e);
      const auto m = system.internal(e);
      for (int q = 0; q < points; ++q) {
        const auto& a = all[static_cast<std::size_t>(e)][static_cast<std::size_t>(q)];
        const auto& b = one[static_cast<std::size_t>(q)];
        expect(a.plastic_strain == b.plastic_strain && a.back_stresses == b.back_stresses &&
                   a.equivalent_plastic_strain == b.equivalent_plastic_strain && a.loading == b.loading,
               "committed(e) equals committed_all()[e]");
        top = std::fmax(top, a.equivalent_plastic_strain);
        put_state(out, a);
      }
      expect(m == modes[static_cast<std::size_t>(e)], "internal(e) equals internal_all()[e]");
      for (double x : m) out.put(x);
    }
    expect(top == system.max_plastic_strain(), "max_plastic_strain is the committed states' largest");
    const std::vector<double> stress = system.stress(u.back());
    out.put(stress);
    // a second system restarted from committed_all / internal_all evaluates
    // the next displacement as the first does
    sparlab_gpu::System restarted(coords, conn, material(static_cast<int>(which)), rule, SPARLAB_FINITE_LOGARITHMIC);
    restarted.set_committed(all);
    restarted.set_internal(modes);
    std::vector<double> next = u.back();
    for (double& x : next) x *= 1.05;
    const auto a = system.evaluate(next, 0.0, true);
    const auto b = restarted.evaluate(next, 0.0, true);
    expect(a.internal == b.internal && a.tangent_values == b.tangent_values, "a restart evaluates as the original");
    // a mismatched restart is refused with the reference's ConfigError
    bool refused = false;
    try {
      restarted.set_committed(std::vector<std::vector<sparlab_gpu::PlasticState>>(static_cast<std::size_t>(ne + 1)));
    } catch (const sparlab_gpu::ConfigError&) {
      refused = true;
    }
    expect(refused, "a history of another element count is a ConfigError");
    std::ofstream of(argv[2], std::ios::binary);
    of.write(reinterpret_cast<const char*>(out.v.data()), static_cast<std::streamsize>(out.v.size() * sizeof(double)));
    std::printf("adapter sequence: %lld elements, %d points, %lld increments, %d failed checks\n",
                static_cast<long long>(ne), points, static_cast<long long>(steps), failures);
    return failures == 0 ? 0 : 1;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "adapter sequence: %s\n", ex.what());
    return 1;
  }
}
