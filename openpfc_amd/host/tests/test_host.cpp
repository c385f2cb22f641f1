// test_host.cpp -- checks of the C++ host layer (driven by tests/test_host_cpp.py).
//
//   test_host cpu                      Time/Simulator/World/modifier semantics
//                                      against a host-only IFFT stand-in
//   test_host gpu-fft                  HipFft overloads, known answers, errors
//   test_host gpu-sim N STEPS OUT.bin  PFCModelHIP driven by Simulator from
//                                      Constant(-0.2) + UniformNoise(1234, +-0.01);
//                                      writes the final field (raw f64) and
//                                      prints its free energy
//   test_host gpu-tung STEPS OUT MFOUT TungstenHIP<double> on the reference's
//                                      CPU-vs-HIP recipe (32^3, spacing 1,
//                                      T = 0.5, dt = 0.01, -0.4 + 0.1 sin(2 pi i/n));
//                                      writes psi and psi_MF (raw f64)
//   test_host seeds DIR                host SeedGrid / RandomSeeds / MovingBC
//                                      (see run_seeds)
//   test_host gpu-callsites N STEPS    the reference's factory and constructor
//                                      signatures (see run_gpu_callsites)
//   test_host gpu-autotune DIR         the autotune cache (see run_gpu_autotune)
#include <openpfc_amd/app.hpp>
#include <openpfc_amd/openpfc_amd.hpp>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>
#include <unistd.h>

using namespace pfc;

static int g_fail = 0;
#define CHECK(cond)                                                                 \
  do {                                                                              \
    if (!(cond)) {                                                                  \
      std::fprintf(stderr, "CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++g_fail;                                                                     \
    }                                                                               \
  } while (0)

template <typename Ex, typename F> static bool throws(F &&f, const char *needle = nullptr) {
  try {
    f();
  } catch (const Ex &e) {
    return !needle || std::strstr(e.what(), needle) != nullptr;
  } catch (...) {
    return false;
  }
  return false;
}

// Host-only IFFT with the one-rank boxes (no transforms): lets the CPU checks
// exercise Model, field::apply and the modifiers without a device.
struct BoxOnlyFft final : fft::IFFT {
  Int3 n;
  explicit BoxOnlyFft(Int3 s) : n(s) {}
  void forward(const fft::RealVector &, fft::ComplexVector &) override {
    throw std::runtime_error("BoxOnlyFft: no transforms");
  }
  void backward(const fft::ComplexVector &, fft::RealVector &) override {
    throw std::runtime_error("BoxOnlyFft: no transforms");
  }
  void reset_fft_time() override {}
  double get_fft_time() const override { return 0; }
  size_t size_inbox() const override { return (size_t)n[0] * n[1] * n[2]; }
  size_t size_outbox() const override { return (size_t)(n[0] / 2 + 1) * n[1] * n[2]; }
  size_t size_workspace() const override { return size_outbox(); }
  size_t get_allocated_memory_bytes() const override { return 0; }
  Box3i get_inbox_bounds() const override { return {{0, 0, 0}, {n[0] - 1, n[1] - 1, n[2] - 1}, n}; }
  Box3i get_outbox_bounds() const override {
    return {{0, 0, 0}, {n[0] / 2, n[1] - 1, n[2] - 1}, {n[0] / 2 + 1, n[1], n[2]}};
  }
};

struct CountingModel final : Model {
  RealField psi;
  int steps = 0;
  std::vector<double> times;
  int prepares = 0, finalizes = 0;
  CountingModel(fft::IFFT &f, const World &w) : Model(f, w) {}
  void initialize(double) override {
    psi.assign(get_fft().size_inbox(), 0.0);
    add_real_field("default", psi);
  }
  void step(double t) override {
    ++steps;
    times.push_back(t);
  }
  void prepare_for_field_modifiers() override { ++prepares; }
  void finalize_after_field_modifiers() override { ++finalizes; }
};

static int run_cpu() {
  // Time: t in [0, 100], dt = 0.5 -> exactly 200 model.step calls
  {
    World w = domain::create(GridSize({8, 4, 2}), PhysicalOrigin({0, 0, 0}), GridSpacing({1, 1, 1}));
    BoxOnlyFft f(w.size);
    CountingModel m(f, w);
    m.initialize(0.5);
    Time t({0.0, 100.0, 0.5}, 5.0);
    Simulator sim(m, t);
    sim.add_initial_conditions(std::make_unique<Constant>(-0.2));
    int saves = 0;
    sim.add_results_writer([&](Model &, const Time &) { ++saves; });
    while (!done(sim)) pfc::step(sim);
    CHECK(m.steps == 200);
    CHECK(t.get_increment() == 200);
    CHECK(std::abs(m.times.front() - 0.5) < 1e-12 && std::abs(m.times.back() - 100.0) < 1e-12);
    CHECK(saves == 21); // t = 0, 5, ..., 100
    CHECK(m.psi[0] == -0.2 && m.psi.back() == -0.2);
    CHECK(m.prepares >= 1 && m.finalizes == 1); // ICs once, no BCs
  }
  // step_block: the same model steps, times and saves as step() calls, in
  // blocks that end at save points; with a boundary condition, one step at
  // a time
  {
    World w = domain::create(GridSize({8, 4, 2}), PhysicalOrigin({0, 0, 0}), GridSpacing({1, 1, 1}));
    BoxOnlyFft f(w.size);
    CountingModel a(f, w), b(f, w);
    a.initialize(0.5);
    b.initialize(0.5);
    Time ta({0.0, 100.0, 0.5}, 5.0), tb({0.0, 100.0, 0.5}, 5.0);
    Simulator sa(a, ta), sb(b, tb);
    std::vector<double> save_a, save_b;
    sa.add_results_writer([&](Model &, const Time &t) { save_a.push_back(t.get_current()); });
    sb.add_results_writer([&](Model &, const Time &t) { save_b.push_back(t.get_current()); });
    while (!done(sa)) pfc::step(sa);
    int calls = 0;
    while (!done(sb)) {
      const int k = sb.step_block(1 << 30);
      CHECK(k >= 1 && k <= 10);
      ++calls;
    }
    CHECK(b.steps == 200 && b.times == a.times && save_b == save_a);
    CHECK(calls == 21); // the first step, then 20 blocks of 10 (saveat 5, dt 0.5)
    CountingModel c(f, w);
    c.initialize(0.5);
    Time tc({0.0, 10.0, 0.5}, 5.0);
    Simulator sc(c, tc);
    sc.add_boundary_conditions(std::make_unique<Constant>(-0.2));
    int one = 0;
    while (!done(sc)) one += sc.step_block(1 << 30) == 1;
    CHECK(c.steps == 20 && one == 20);
  }
  // Time validation (time.hpp:366-395)
  CHECK(throws<std::invalid_argument>([] { Time t({-1.0, 1.0, 0.1}, 0.0); }));
  CHECK(throws<std::invalid_argument>([] { Time t({1.0, 1.0, 0.1}, 0.0); }));
  CHECK(throws<std::invalid_argument>([] { Time t({0.0, 1.0, 0.1}, 2.0); }));
  // World coordinates and k-space helpers
  {
    World w = domain::create_centered(GridSize({64, 64, 64}), GridSpacing({M_PI / 4, M_PI / 4, M_PI / 4}));
    CHECK(std::abs(w.origin[0] + 32 * M_PI / 4) < 1e-12);
    Real3 x = domain::to_coords(w, {32, 0, 63});
    CHECK(x[0] == 0.0 && x[1] == w.origin[1] && std::abs(x[2] - (w.origin[2] + 63 * M_PI / 4)) < 1e-12);
    auto f = fft::kspace::k_frequency_scaling(w);
    CHECK(std::abs(f[0] - 0.125) < 1e-15);
    CHECK(fft::kspace::k_component(8, 64, f[0]) == 1.0);
    CHECK(fft::kspace::k_component(63, 64, f[0]) == -0.125);
    CHECK(fft::kspace::k_laplacian_value(1, 2, 3) == -14.0);
  }
  // SingleSeed: rho + 12 amp at the origin, 0 outside
  {
    World w = domain::create(GridSize({16, 16, 16}), PhysicalOrigin({-8, -8, -8}), GridSpacing({1, 1, 1}));
    BoxOnlyFft f(w.size);
    CountingModel m(f, w);
    m.initialize(1.0);
    SingleSeed s;
    s.set_density(-0.1);
    s.set_amplitude(0.05);
    s.set_radius(4.0);
    s.apply(m, 0.0);
    const size_t c = (8 * 16 + 8) * 16 + 8;
    CHECK(std::abs(m.psi[c] - (-0.1 + 12 * 0.05)) < 1e-15);
    CHECK(m.psi[0] == 0.0);
    // field size mismatch is rejected
    RealField bad(10);
    CHECK(throws<std::invalid_argument>(
        [&] { field::apply(bad, w, f, [](const Real3 &) { return 0.0; }); }, "inbox"));
    CHECK(throws<std::invalid_argument>([&] { m.get_real_field("nope"); }));
  }
  // Binary field IO on host buffers (test_binary_writer.cpp recipes): two
  // x bricks of a {4,2,2} array, round trip, layout, FileReader, errors
  {
    const std::string tmpl = std::string("/tmp/openpfc_amd_host_io_") +
                             std::to_string((long)getpid()) + "_%d.bin";
    const std::array<int, 3> g{4, 2, 2}, l{2, 2, 2};
    std::vector<double> whole(16);
    for (int r = 0; r < 2; ++r) {
      RealField d(8);
      for (int i = 0; i < 8; ++i) d[i] = i + 1000.0 * r + 0.25 * (2 * r);
      BinaryWriter w(tmpl);
      w.set_domain(g, l, {2 * r, 0, 0});
      if (r == 0) w.write(0, d);
      else
        CHECK(pfc_hip_field_write_host(utils::format_with_number(tmpl, 0).c_str(), d.data(),
                                       PFC_HIP_F64, PFC_HIP_IO_REAL, g.data(), l.data(),
                                       std::array<int, 3>{2, 0, 0}.data(), 0) == 0);
      for (int i = 0; i < 8; ++i) whole[(i / 2) * 4 + 2 * r + (i % 2)] = d[i];
    }
    const std::string path = utils::format_with_number(tmpl, 0);
    RealField all(16);
    BinaryReader rd;
    rd.set_domain(g, g, {0, 0, 0});
    rd.read(path, all);
    CHECK(all == whole);
    // FileReader ("from_file") into a model's field
    World w = domain::create(GridSize({4, 2, 2}), PhysicalOrigin({0, 0, 0}), GridSpacing({1, 1, 1}));
    BoxOnlyFft f(w.size);
    CountingModel m(f, w);
    m.initialize(1.0);
    FileReader fr(path);
    fr.apply(m, 0.0);
    CHECK(m.psi == whole);
    FileReader missing("/nonexistent/dir/none.bin");
    CHECK(throws<std::runtime_error>([&] { missing.apply(m, 0.0); }, "FileReader failed"));
    RealField wrong(15);
    CHECK(throws<std::runtime_error>([&] { rd.read(path, wrong); }, "buffer size mismatch"));
    BinaryWriter nodomain(tmpl);
    CHECK(throws<std::runtime_error>([&] { nodomain.write(0, whole); }, "set_domain"));
    CHECK(throws<std::invalid_argument>([&] { rd.set_domain({8, 8, 8}, {8, 8, 8}, {4, 0, 0}); }));
    CHECK(utils::format_with_number("frame_%04d.vti", 3) == "frame_0003.vti");
    CHECK(throws<std::invalid_argument>([] { utils::format_with_number("a_%s", 1); }));
    std::remove(path.c_str());
  }
  // backend names (backend_from_string.hpp)
  CHECK(fft::backend_from_string("hip") == fft::Backend::HIP);
  CHECK(fft::backend_from_string("rocm") == fft::Backend::HIP);
  CHECK(throws<std::invalid_argument>([] { fft::backend_from_string("opencl"); }));
  // Decomposition / FFTLayout / factories with the reference's signatures
  // (decomposition.hpp, fft_layout.hpp, fft_hip.hpp:65-78, fft.hpp:52-57):
  // the refusals happen before any device work
  {
    World w = world::create(GridSize({16, 8, 12}), PhysicalOrigin({0, 0, 0}),
                            GridSpacing({1, 1, 1}));
    auto d1 = decomposition::create(w, 1);
    CHECK(decomposition::get_num_domains(d1) == 1);
    CHECK(decomposition::get_local_box(d1, 0).size == Int3({16, 8, 12}));
    auto d4 = decomposition::create(w, 4);
    CHECK(decomposition::get_grid(d4) == (std::array<int, 3>{1, 1, 4}));
    CHECK(decomposition::get_local_box(d4, 3).low[2] == 9 &&
          decomposition::get_local_box(d4, 3).high[2] == 11);
    auto l1 = fft::layout::create(d1, 0);
    CHECK(fft::layout::get_complex_box(l1, 0).size == Int3({9, 8, 12}));
    CHECK(fft::layout::get_real_box(l1, 0).high == Int3({15, 7, 11}));
    CHECK(fft::layout::get_r2c_direction(l1) == 0);
    auto l4 = fft::layout::create(d4, 0);
    CHECK(fft::layout::get_complex_box(l4, 2).low[2] == 6 &&
          fft::layout::get_complex_box(l4, 2).size[0] == 9);
    CHECK(throws<std::invalid_argument>([&] { (void)fft::layout::create(d1, 1); }));
    CHECK(throws<std::logic_error>([&] { (void)fft::create_hip(d4, 0, MPI_COMM_WORLD); },
                                   "one part per GPU"));
    CHECK(throws<std::invalid_argument>([&] { (void)fft::create_hip(d1, 1); }));
    CHECK(throws<std::runtime_error>([&] {
      (void)fft::create_with_backend(l1, 0, fft::default_options(), fft::Backend::FFTW,
                                     MPI_COMM_WORLD);
    }, "Unsupported FFT backend"));
    CHECK(throws<std::logic_error>([&] {
      (void)fft::create_with_backend(d4, 0, fft::Backend::HIP);
    }, "one part per GPU"));
    // a Model over an IFFT that is not a HipFft is refused by the device models
    BoxOnlyFft bf(w.size);
    CHECK(throws<std::invalid_argument>([&] { PFCModelHIP m(bf, w, -0.25); }, "HipFft"));
    CHECK(throws<std::invalid_argument>([&] { TungstenHIP<double> m(bf, w, MPI_COMM_WORLD); },
                                        "HipFft"));
    CountingModel cm(bf, w);
    Time t({0.0, 1.0, 0.5}, 0.0);
    Simulator sim(cm, t, MPI_COMM_WORLD);
    CHECK(sim.mpi_comm() == MPI_COMM_WORLD && cm.mpi_comm() == MPI_COMM_WORLD);
  }
  return g_fail;
}

// Reference call sites on the device: the FFT from create_hip(decomposition,
// rank, MPI_COMM_WORLD) and from create_with_backend(layout, rank, options,
// Backend::HIP, MPI_COMM_WORLD), the model and the Simulator with their
// communicator arguments; both runs must give the field of the World-based
// construction (gpu-sim) bit for bit.
static int run_gpu_callsites(int n, int steps) {
  const double dx = M_PI / 4, dt = 0.5;
  World w = domain::create_centered(GridSize({n, n, n}), GridSpacing({dx, dx, dx}));
  auto run = [&](fft::IFFT &f) {
    PFCModelHIP model(f, w, -0.25, MPI_COMM_WORLD);
    initialize(model, dt);
    Time time({0.0, steps * dt, dt}, 0.0);
    Simulator sim(model, time, MPI_COMM_WORLD);
    auto c = std::make_unique<Constant>(-0.2);
    c->set_field_name("density");
    sim.add_initial_conditions(std::move(c));
    auto noise = std::make_unique<UniformNoise>(1234, -0.01, 0.01);
    noise->set_field_name("density");
    sim.add_initial_conditions(std::move(noise));
    while (!done(sim)) pfc::step(sim);
    model.prepare_for_field_modifiers();
    return get_real_field(model, "density");
  };
  auto decomp = decomposition::create(w, 1);
  fft::HipFft f0(w);
  const RealField base = run(f0);
  auto f1 = fft::create_hip(decomp, 0, MPI_COMM_WORLD);
  CHECK(run(f1) == base);
  auto layout = fft::layout::create(decomp, 0);
  auto f2 = fft::create_with_backend(layout, 0, fft::default_options(), fft::Backend::HIP,
                                     MPI_COMM_WORLD);
  CHECK(run(*f2) == base);
  auto f3 = fft::create_with_backend(decomp, 0, fft::Backend::HIP);
  CHECK(run(*f3) == base);
  std::printf("callsites %d^3 x %d steps: %zu values\n", n, steps, base.size());
  return g_fail;
}

static std::vector<std::complex<double>> dft_r2c(const std::vector<double> &a, Int3 n) {
  const int w = n[0] / 2 + 1;
  std::vector<std::complex<double>> out((size_t)w * n[1] * n[2]);
  for (int kz = 0; kz < n[2]; ++kz)
    for (int ky = 0; ky < n[1]; ++ky)
      for (int kx = 0; kx < w; ++kx) {
        std::complex<long double> s = 0;
        for (int z = 0; z < n[2]; ++z)
          for (int y = 0; y < n[1]; ++y)
            for (int x = 0; x < n[0]; ++x) {
              long double ang = -2.0L * M_PIl *
                                ((long double)(kx * x % n[0]) / n[0] +
                                 (long double)(ky * y % n[1]) / n[1] +
                                 (long double)(kz * z % n[2]) / n[2]);
              s += (long double)a[((size_t)z * n[1] + y) * n[0] + x] *
                   std::complex<long double>(cosl(ang), sinl(ang));
            }
        out[((size_t)kz * n[1] + ky) * w + kx] = {(double)s.real(), (double)s.imag()};
      }
  return out;
}

static int run_gpu_fft() {
  World w = domain::create(GridSize({8, 4, 2}), PhysicalOrigin({0, 0, 0}), GridSpacing({1, 1, 1}));
  fft::HipFft f(w);
  CHECK(f.size_inbox() == 64 && f.size_outbox() == 5 * 4 * 2 && f.size_workspace() == 40);
  Box3i ob = f.get_outbox_bounds();
  CHECK(ob.high[0] == 4 && ob.size[0] == 5);
  std::vector<double> a(64);
  for (size_t i = 0; i < a.size(); ++i) a[i] = std::sin(0.37 * i) + 0.1 * (i % 5);
  auto ref = dft_r2c(a, w.size);
  // host-vector overloads (staged)
  std::vector<std::complex<double>> X(f.size_outbox());
  f.forward(a, X);
  double err = 0, mx = 0;
  for (size_t i = 0; i < X.size(); ++i) {
    err = std::max(err, std::abs(X[i] - ref[i]));
    mx = std::max(mx, std::abs(ref[i]));
  }
  CHECK(err <= 1e-13 * mx);
  std::vector<double> back(64);
  f.backward(X, back);
  for (size_t i = 0; i < a.size(); ++i) CHECK(std::abs(back[i] - a[i]) <= 1e-14);
  CHECK(f.get_fft_time() > 0);
  // DataBuffer overloads
  core::DataBuffer<backend::HipTag, double> din(64), dout(64);
  core::DataBuffer<backend::HipTag, std::complex<double>> dX(40);
  din.copy_from_host(a);
  f.forward(din, dX);
  auto hX = dX.to_host();
  for (size_t i = 0; i < hX.size(); ++i) CHECK(std::abs(hX[i] - X[i]) <= 1e-15 * mx);
  f.backward(dX, dout);
  auto hb = dout.to_host();
  for (size_t i = 0; i < hb.size(); ++i) CHECK(std::abs(hb[i] - a[i]) <= 1e-14);
  // size and precision errors (fft_heffte_backend.hpp:53-61)
  std::vector<double> small(63);
  CHECK(throws<std::invalid_argument>([&] { f.forward(small, X); },
                                      "FFT_Impl::forward: real buffer size 63 != size_inbox() 64"));
  std::vector<std::complex<double>> Xs(39);
  CHECK(throws<std::invalid_argument>([&] { f.forward(a, Xs); }, "size_outbox"));
  CHECK(throws<std::invalid_argument>([&] { f.backward(Xs, back); }, "size_outbox"));
  core::DataBuffer<backend::HipTag, float> fin(64);
  core::DataBuffer<backend::HipTag, std::complex<float>> fX(40);
  CHECK(throws<std::runtime_error>([&] { f.forward(fin, fX); }, "precision"));
  // reference KAT (test_fft.cpp:59-70): 2x1x1 backward of {1, 2} -> 1.5
  {
    World w2 = domain::create(GridSize({2, 1, 1}), PhysicalOrigin({0, 0, 0}), GridSpacing({1, 1, 1}));
    fft::HipFft f2(w2);
    std::vector<std::complex<double>> in = {{1, 0}, {2, 0}};
    std::vector<double> out(2);
    f2.backward(in, out);
    CHECK(std::abs(out[0] - 1.5) <= 0.01);
  }
  // unsupported extents raise
  CHECK(throws<std::runtime_error>([] {
    fft::HipFft bad(domain::create(GridSize({6, 6, 6}), PhysicalOrigin({0, 0, 0}),
                                   GridSpacing({1, 1, 1})));
  }, "unsupported"));
  // float plan through DataBuffer<float>
  {
    fft::HipFft ff(w, PFC_HIP_F32);
    std::vector<float> af(a.begin(), a.end());
    core::DataBuffer<backend::HipTag, float> fa(64);
    fa.copy_from_host(af);
    ff.forward(fa, fX);
    auto h = fX.to_host();
    for (size_t i = 0; i < h.size(); ++i)
      CHECK(std::abs(std::complex<double>(h[i]) - ref[i]) <= 2e-6 * mx);
  }
  return g_fail;
}

static int run_gpu_sim(int n, int steps, const char *out) {
  const double dx = M_PI / 4, dt = 0.5;
  World w = domain::create_centered(GridSize({n, n, n}), GridSpacing({dx, dx, dx}));
  fft::HipFft f(w);
  PFCModelHIP model(f, w, -0.25);
  initialize(model, dt);
  Time time({0.0, steps * dt, dt}, 0.0);
  Simulator sim(model, time);
  auto c = std::make_unique<Constant>(-0.2);
  c->set_field_name("density");
  sim.add_initial_conditions(std::move(c));
  auto noise = std::make_unique<UniformNoise>(1234, -0.01, 0.01);
  noise->set_field_name("density");
  sim.add_initial_conditions(std::move(noise));
  while (!done(sim)) pfc::step(sim);
  model.prepare_for_field_modifiers();
  const RealField &psi = get_real_field(model, "density");
  std::ofstream o(out, std::ios::binary);
  o.write(reinterpret_cast<const char *>(psi.data()), (std::streamsize)(psi.size() * sizeof(double)));
  std::printf("steps %d increment %d\n", steps, time.get_increment());
  std::printf("free_energy %.17g\n", model.free_energy());
  CHECK(time.get_increment() == steps);
  return g_fail;
}

// apps/tungsten/tests/test_tungsten_cpu_vs_hip.cpp:30-90 through the C++
// drop-in: TungstenHIP<double> over a HipFft, psi set through the host field.
static int run_gpu_tung(int steps, const char *out, const char *mfout) {
  World w = domain::create(GridSize({32, 32, 32}), PhysicalOrigin({0.0, 0.0, 0.0}),
                           GridSpacing({1.0, 1.0, 1.0}));
  fft::HipFft f(w);
  TungstenHIP<double> model(f, w);
  model.params.set_n0(-0.4);
  model.params.set_T(0.5);
  initialize(model, 0.01);
  RealField &psi = get_real_field(model, "psi");
  for (size_t i = 0; i < psi.size(); ++i)
    psi[i] = -0.4 + 0.1 * std::sin(2.0 * M_PI * (double)i / (double)psi.size());
  model.finalize_after_field_modifiers();
  for (int i = 0; i < steps; ++i) model.step(0.0);
  const RealField &res = model.get_psi_for_writer();
  const auto mf = model.get_psiMF().to_host();
  std::ofstream o(out, std::ios::binary);
  o.write(reinterpret_cast<const char *>(res.data()), (std::streamsize)(res.size() * sizeof(double)));
  std::ofstream m(mfout, std::ios::binary);
  m.write(reinterpret_cast<const char *>(mf.data()), (std::streamsize)(mf.size() * sizeof(double)));
  CHECK(model.get_allocated_memory_bytes() > 0);
  // a float plan does not fit TungstenHIP<double>
  fft::HipFft f32(w, PFC_HIP_F32);
  CHECK(throws<std::invalid_argument>([&] { TungstenHIP<double> bad(f32, w); }));
  return g_fail;
}

// Settings files: canonical JSON of a parsed .json/.toml file (checked
// against Python's json and tomli by the tests), and World/Time wiring.
static int run_parse(const char *path) {
  std::printf("%s\n", ui::load_settings_file(path).dump().c_str());
  return 0;
}

static int run_config(const char *path) {
  const json::Value s = ui::load_settings_file(path);
  World w = ui::world_from_json(s);
  Time t = ui::time_from_json(s);
  std::printf("world %d %d %d spacing %.17g %.17g %.17g origin %.17g %.17g %.17g\n", w.size[0],
              w.size[1], w.size[2], w.spacing[0], w.spacing[1], w.spacing[2], w.origin[0],
              w.origin[1], w.origin[2]);
  std::printf("time %.17g %.17g %.17g %.17g\n", t.get_t0(), t.get_t1(), t.get_dt(),
              t.get_saveat());
  if (const json::Value *ics = s.find("initial_conditions"))
    for (const json::Value &j : ics->items()) (void)ui::field_modifier_from_json(j);
  if (const json::Value *bcs = s.find("boundary_conditions"))
    for (const json::Value &j : bcs->items()) (void)ui::field_modifier_from_json(j);
  ui::check_backend(s);
  return 0;
}

// Host paths of the seed modifiers and MovingBC, for the oracle comparison
// in tests/test_seeds_host.py.  Writes into DIR (raw little-endian f64):
//   grid_rows.bin    SeedGrid rows (24 doubles each) on 128x64x64, dx 1.1107...
//   random_rows.bin  RandomSeeds rows (150 x 24)
//   draws.bin        the first 16 std::uniform_real_distribution(-3, 5) draws
//                    of std::mt19937_64(42)
//   grid_field.bin   SeedGrid field (host field::apply path) on that grid
//   random_field.bin RandomSeeds field on 64^3, spacing 4, origin -128
//   moving_%d.bin    a field after MovingBC call d = 0, 1, 2 (xline and xpos
//                    printed), from the SeedGrid field shifted each call
static int run_seeds(const std::string &dir) {
  auto wr = [&](const std::string &name, const std::vector<double> &v) {
    std::ofstream f(dir + "/" + name, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
  };
  const double dx = 1.1107207345395915;
  World w = domain::create(GridSize({128, 64, 64}), PhysicalOrigin({0, 0, 0}),
                           GridSpacing({dx, dx, dx}));
  SeedGrid g;
  g.set_Ny(2);
  g.set_Nz(2);
  g.set_X0(30.0);
  g.set_radius(25.0);
  g.set_amplitude(0.215936);
  g.set_density(-0.047);
  std::vector<double> rows;
  for (const Seed &sd : g.seeds(w)) {
    double r[24];
    sd.to_row(r);
    rows.insert(rows.end(), r, r + 24);
  }
  wr("grid_rows.bin", rows);
  RandomSeeds rs;
  rs.set_amplitude(0.215936);
  rs.set_density(-0.047);
  rows.clear();
  for (const Seed &sd : rs.seeds()) {
    double r[24];
    sd.to_row(r);
    rows.insert(rows.end(), r, r + 24);
  }
  wr("random_rows.bin", rows);
  {
    std::mt19937_64 re(42);
    std::uniform_real_distribution<double> u(-3.0, 5.0);
    std::vector<double> d(16);
    for (double &x : d) x = u(re);
    wr("draws.bin", d);
  }
  {
    BoxOnlyFft f(w.size);
    CountingModel m(f, w);
    m.initialize(1.0);
    m.psi.assign(m.psi.size(), -0.1);
    g.set_field_name("default");
    g.apply(m, 0.0);
    wr("grid_field.bin", m.psi);
    MovingBC b(-0.464, -0.1);
    b.set_xwidth(15.0);
    b.set_alpha(1.0);
    b.set_disp(40.0);
    b.set_xpos(0.0);
    b.set_field_name("default");
    RealField base = m.psi;
    for (int call = 0; call < 3; ++call) {
      // move the seeded columns 9 cells along x per call
      for (size_t r = 0; r < base.size() / 128; ++r)
        for (int i = 0; i < 128; ++i) m.psi[r * 128 + i] = base[r * 128 + (i + 128 - 9 * call) % 128];
      b.apply(m, 0.0);
      std::printf("moving %d idx %d xpos %.17g\n", call, b.get_index(), b.get_xpos());
      wr("moving_" + std::to_string(call) + ".bin", m.psi);
    }
  }
  {
    World w2 = domain::create(GridSize({64, 64, 64}), PhysicalOrigin({-128, -128, -128}),
                              GridSpacing({4, 4, 4}));
    BoxOnlyFft f(w2.size);
    CountingModel m(f, w2);
    m.initialize(1.0);
    rs.set_field_name("default");
    rs.apply(m, 0.0);
    wr("random_field.bin", m.psi);
  }
  return g_fail;
}

// The autotune cache of the C++ drop-in (autotune.hpp) in directory DIR:
// the first model of a size tunes and stores its choice, the next one loads
// it; cache_only refuses a size without an entry; DISABLE keeps defaults;
// an entry naming unknown variants counts as missing.
static int run_gpu_autotune(const std::string &dir) {
  setenv("OPENPFC_GPU_AUTOTUNE_CACHE_DIR", dir.c_str(), 1);
  unsetenv("OPENPFC_GPU_AUTOTUNE_DISABLE");
  setenv("OPENPFC_GPU_AUTOTUNE_MODE", "auto", 1);
  const double dx = M_PI / 4;
  World w = domain::create_centered(GridSize({64, 32, 32}), GridSpacing({dx, dx, dx}));
  fft::HipFft f(w);
  PFCModelHIP a(f, w, -0.25);
  a.initialize(0.5);
  CHECK(a.autotune_outcome() == gpu::TuneOutcome::tuned);
  const std::string path = gpu::cache_file();
  CHECK(std::filesystem::exists(path));
  const json::Value cache = gpu::load_cache(path);
  const json::Value *e = cache.find("entries");
  CHECK(e && e->find(gpu::step_key("pfc", PFC_HIP_F64, w.size, false)));
  const json::Value chosen = gpu::current_entry(f.plan());
  PFCModelHIP b(f, w, -0.25);
  b.initialize(0.5);
  CHECK(b.autotune_outcome() == gpu::TuneOutcome::cached);
  CHECK(gpu::current_entry(f.plan()).dump() == chosen.dump());
  World w2 = domain::create_centered(GridSize({32, 32, 32}), GridSpacing({dx, dx, dx}));
  fft::HipFft f2(w2);
  setenv("OPENPFC_GPU_AUTOTUNE_MODE", "cache_only", 1);
  CHECK(throws<std::runtime_error>([&] {
    PFCModelHIP c(f2, w2, -0.25);
    c.initialize(0.5);
  }, "No cached configuration"));
  setenv("OPENPFC_GPU_AUTOTUNE_DISABLE", "1", 1);
  PFCModelHIP d(f2, w2, -0.25);
  d.initialize(0.5);
  CHECK(d.autotune_outcome() == gpu::TuneOutcome::defaults);
  unsetenv("OPENPFC_GPU_AUTOTUNE_DISABLE");
  // an entry with a variant name the library does not offer is a miss
  json::Value bad = gpu::load_cache(path);
  bad["entries"][gpu::step_key("pfc", PFC_HIP_F64, w.size, false)]["4"] =
      json::Value::string("no-such-variant");
  gpu::store_cache(path, bad);
  setenv("OPENPFC_GPU_AUTOTUNE_MODE", "auto", 1);
  PFCModelHIP g(f, w, -0.25);
  g.initialize(0.5);
  CHECK(g.autotune_outcome() == gpu::TuneOutcome::tuned);
  std::printf("autotune cache %s\n", path.c_str());
  return g_fail;
}

int main(int argc, char **argv) {
  const std::string mode = argc > 1 ? argv[1] : "cpu";
  int rc = 0;
  try {
    if (mode == "cpu") rc = run_cpu();
    else if (mode == "gpu-fft") rc = run_gpu_fft();
    else if (mode == "gpu-sim" && argc == 5) rc = run_gpu_sim(std::atoi(argv[2]), std::atoi(argv[3]), argv[4]);
    else if (mode == "gpu-tung" && argc == 5) rc = run_gpu_tung(std::atoi(argv[2]), argv[3], argv[4]);
    else if (mode == "parse" && argc == 3) return run_parse(argv[2]);
    else if (mode == "config" && argc == 3) rc = run_config(argv[2]);
    else if (mode == "seeds" && argc == 3) rc = run_seeds(argv[2]);
    else if (mode == "gpu-autotune" && argc == 3) rc = run_gpu_autotune(argv[2]);
    else if (mode == "gpu-callsites" && argc == 4)
      rc = run_gpu_callsites(std::atoi(argv[2]), std::atoi(argv[3]));
    else {
      std::fprintf(stderr, "usage: test_host cpu | gpu-fft | gpu-sim N STEPS OUT | "
                           "gpu-tung STEPS OUT MFOUT | parse FILE | config FILE | seeds DIR | "
                           "gpu-callsites N STEPS | gpu-autotune DIR\n");
      return 2;
    }
  } catch (const std::exception &e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 3;
  }
  std::printf("%s: %s (%d failed checks)\n", mode.c_str(), rc ? "FAIL" : "ok", rc);
  return rc ? 1 : 0;
}
